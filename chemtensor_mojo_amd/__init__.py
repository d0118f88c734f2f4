"""chemtensor_mojo_amd: the two-site DMRG hot path of chemtensor_mojo on MI355X (gfx950).

Everything numeric runs in libdmrg_hip.so (hand-written HIP kernels for CDNA4:
MFMA GEMMs in FP64 / FP32 / 3xBF16 / 6xBF16, MPO mixing, device Lanczos, rocSOLVER
SVD, C++ sweep driver).  The Python modules mirror the reference's Mojo call
surface (src/m_tensor, src/state, src/algorithms) on top of the C ABI.
"""
from ._lib import PREC_BF16X3, PREC_BF16X6, PREC_FP32, PREC_FP64, SVD_GESVD, SVD_GRAM
from .dense_tensor import (DenseTensor, DeviceContext, axpy_in_place, compute_row_major_strides, create_dense_tensor,
                           create_dense_tensor_from_data, create_dense_tensor_uninitialized, dense_tensor_dot,
                           dot_product)
from .dmrg import (DMRGParams, DMRGRun, DMRGStats, HeffOperator, apply_two_site_heff, build_two_site_theta, dense_tensor_svd_trunc,
                   dmrg_two_site, lanczos_two_site_optimize, split_two_site_theta_svd)
from .state import (MatrixProductOperator, MatrixProductState, MPOSite, MPSSite, build_right_environments,
                    create_product_mps, create_random_mps_c_compatible, expectation_value, mpo_from_tensors,
                    mps_from_tensors, random_mps_c_compatible_tensors, update_left_environment,
                    update_right_environment)
from .results import (DMRGJsonParams, append_bench_record, bond_entropies, dmrg_results_to_json_string,
                      load_dmrg_reference, mps_norm, save_dmrg_results_to_json)
from . import hamiltonians

__version__ = "0.1.0"
