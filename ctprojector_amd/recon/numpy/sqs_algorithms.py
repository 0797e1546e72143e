"""The reference's src/ct_projector/recon/numpy/sqs_algorithms.py: the shared
implementation in ctprojector_amd.recon.sqs_algorithms."""
from ..sqs_algorithms import nesterov_acceleration, sqs_one_step, sqs_tv_one_step  # noqa: F401
