"""CPU: the C-ABI library builds for gfx950, loads, and exports exactly the
symbols include/ctprojector.h declares (no compute calls without a GPU); the
error paths return codes instead of aborting."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'ctprojector.h')


def declared():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(?:int|const char\*)\s+(\w+)\s*\(', text)))


def exported(path):
    import subprocess
    out = subprocess.run(['nm', '-D', '--defined-only', path], stdout=subprocess.PIPE, text=True, check=True).stdout
    return sorted(set(line.split()[-1] for line in out.splitlines() if ' T ' in line))


def test_header_lists_the_reference_hot_path():
    names = declared()
    # the 23 hot-path entry points of the reference's libprojector.so
    for n in ['SetDevice', 'cSiddonConeProjectionArbitrary', 'cupySiddonConeProjectionArbitrary',
              'cSiddonConeBackprojectionArbitrary', 'cupySiddonConeBackprojectionArbitrary',
              'cupyDistanceDrivenFanProjection', 'cupyDistanceDrivenParallelProjection',
              'cDistanceDrivenFanProjection', 'cDistanceDrivenParallelProjection',
              'cupyDistanceDrivenFanBackprojection', 'cupyDistanceDrivenParallelBackprojection',
              'cDistanceDrivenFanBackprojection', 'cDistanceDrivenParallelBackprojection',
              'cDistanceDrivenTomoProjection', 'cupyDistanceDrivenTomoProjection',
              'cDistanceDrivenTomoBackprojection', 'cupyDistanceDrivenTomoBackprojection',
              'cupyfbpFanFilter', 'cfbpFanFilter', 'cupyfbpParallelFilter', 'cfbpParallelFilter',
              'cupyfbpFanBackprojection', 'cfbpFanBackprojection']:
        assert n in names, n


def test_library_exports_every_declared_symbol():
    from ctprojector_amd import _lib
    assert os.path.exists(_lib.LIB_PATH), 'run python -m ctprojector_amd.build'
    ex = exported(_lib.LIB_PATH)
    missing = [n for n in declared() if n not in ex]
    assert not missing, missing
    extra = [n for n in ex if n not in declared()]
    assert not extra, 'exported but undeclared: %s' % extra
    assert sorted(_lib.SYMBOLS) == declared()


def test_library_loads_and_fails_cleanly_without_gpu():
    from ctprojector_amd import _lib
    lib = _lib.module()
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip('a GPU is present')
    except ImportError:
        pass
    # no device: an error code, not an abort
    assert lib.SetDevice(ctypes.c_int(0)) != 0
    # invalid arguments are rejected before any device work
    err = lib.cupySiddonConeProjectionArbitrary(
        None, None, None, None, None, None, ctypes.c_size_t(1), ctypes.c_size_t(4), ctypes.c_size_t(4),
        ctypes.c_size_t(4), ctypes.c_float(1), ctypes.c_float(1), ctypes.c_float(1), ctypes.c_float(0),
        ctypes.c_float(0), ctypes.c_float(0), ctypes.c_size_t(4), ctypes.c_size_t(4), ctypes.c_size_t(2),
        ctypes.c_float(1), ctypes.c_float(1), ctypes.c_float(0), ctypes.c_float(0))
    assert err != 0
    assert 'null pointer' in _lib.last_error_message()


def test_tuning_keys():
    from ctprojector_amd import _lib
    _lib.set_tuning('cone_bp_zrun', 2)
    assert _lib.get_tuning('cone_bp_zrun') == 2
    _lib.set_tuning('cone_bp_zrun', 0)
    with pytest.raises(KeyError):
        _lib.set_tuning('no_such_knob', 1)


def test_empty_sizes_are_a_no_op():
    from ctprojector_amd import _lib
    lib = _lib.module()
    err = lib.cfbpFanFilter(None, None, ctypes.c_int(0), ctypes.c_size_t(8), ctypes.c_size_t(1), ctypes.c_size_t(3),
                            ctypes.c_float(0.01), ctypes.c_float(1), ctypes.c_float(0), ctypes.c_float(0),
                            ctypes.c_float(1000), ctypes.c_float(500), ctypes.c_int(2))
    assert err == 0


def test_dd_batch_rejects_bad_jobs_before_device_work():
    """ctpDDBatchCreate validates its job list on the host: no jobs, an empty
    job and missing arrays are errors (codes, not aborts), and no plan is
    returned."""
    from ctprojector_amd import _lib
    from ctprojector_amd.batched import _DJob
    lib = _lib.module()
    plan = ctypes.c_void_p(1)
    assert lib.ctpDDBatchCreate(ctypes.byref(plan), None, ctypes.c_int(0), ctypes.c_int(0)) != 0
    assert 'no jobs' in _lib.last_error_message() and not plan.value
    angles = (ctypes.c_float * 4)(0.0, 0.5, 1.0, 1.5)
    fake = ctypes.c_void_p(4096).value    # never dereferenced: validation fails first

    def job(**kw):
        a = dict(fan=1, typeProjector=0, nBatches=1, nx=8, ny=8, nz=1, dx=1, dy=1, dz=1, cx=0, cy=0, cz=0, nu=12,
                 nv=1, nview=4, da=0.01, dv=1, off_a=0, off_v=0, dsd=200, dso=120,
                 deg=ctypes.cast(angles, ctypes.c_void_p).value, img=fake, prj=fake)
        a.update(kw)
        return _DJob(**a)
    for bad, msg in [(dict(nx=0), 'empty job'), (dict(deg=None), 'null pointer'), (dict(prj=None), 'null pointer')]:
        jobs = (_DJob * 2)(job(), job(**bad))
        plan = ctypes.c_void_p(1)
        assert lib.ctpDDBatchCreate(ctypes.byref(plan), jobs, ctypes.c_int(2), ctypes.c_int(0)) != 0, bad
        assert msg in _lib.last_error_message(), (bad, _lib.last_error_message())
        assert not plan.value


def test_filter_batch_rejects_bad_jobs_before_device_work():
    """ctpFilterBatchCreate validates its job list on the host: no jobs, an
    empty job, a detector row wider than 5000 columns, a zero pitch and
    missing arrays are errors (codes, not aborts), and no plan is returned."""
    from ctprojector_amd import _lib
    from ctprojector_amd.batched import _FJob
    lib = _lib.module()
    plan = ctypes.c_void_p(1)
    assert lib.ctpFilterBatchCreate(ctypes.byref(plan), None, ctypes.c_int(0)) != 0
    assert 'no jobs' in _lib.last_error_message() and not plan.value
    fake = ctypes.c_void_p(4096).value    # never dereferenced: validation fails first

    def job(**kw):
        a = dict(fan=1, typeFilter=0, nBatches=1, nu=12, nv=1, nview=4, da=0.01, off_a=0, dsd=200, dso=120,
                 prj=fake, fprj=fake)
        a.update(kw)
        return _FJob(**a)
    for bad, msg in [(dict(nview=0), 'empty job'), (dict(nu=5001), '5000'), (dict(da=0.0), 'zero detector pitch'),
                     (dict(dso=0.0), 'zero dso'), (dict(prj=None), 'null pointer'), (dict(fprj=None), 'null pointer')]:
        jobs = (_FJob * 2)(job(), job(**bad))
        plan = ctypes.c_void_p(1)
        assert lib.ctpFilterBatchCreate(ctypes.byref(plan), jobs, ctypes.c_int(2)) != 0, bad
        assert msg in _lib.last_error_message(), (bad, _lib.last_error_message())
        assert not plan.value
