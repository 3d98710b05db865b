"""PySCF-intor drop-ins backed by the HIP Rys kernels."""
from .int2e_gpu import compute_int2e_gpu
from .int2e_ip1_gpu import compute_int2e_ip1_gpu
