"""CSR SpMV / SpMM operators (the reference's sparse_operations names for this path)."""

from ._common import SUPPORTED_INDEX_DTYPES, SUPPORTED_VALUE_DTYPES
from .spmv_csr import (
    PreparedCsrSpmv,
    flagsparse_spmv_coo_tocsr,
    flagsparse_spmv_csr,
    prepare_spmv_coo_tocsr,
    prepare_spmv_csr,
)
from .spmm_csr import (
    PreparedCsrSpmmOpt,
    benchmark_spmm_opt_case,
    flagsparse_spmm_csr,
    flagsparse_spmm_csr_opt,
    prepare_spmm_csr_opt,
)
from .spmm_csr_opt_alg2 import (
    PreparedCsrSpmmOptAlg2,
    benchmark_spmm_opt_alg2_case,
    flagsparse_spmm_csr_opt_alg2,
    prepare_spmm_csr_opt_alg2,
)
from .benchmarks import benchmark_spmm_case, benchmark_spmv_case, comprehensive_spmm_test
from .spmv_coo import PreparedCoo, flagsparse_spmv_coo, prepare_spmv_coo
from .spmm_coo import flagsparse_spmm_coo
from .sddmm_csr import (SDDMMPrepared, benchmark_sddmm_case, flagsparse_sddmm_csr,
                         prepare_sddmm_csr)

__all__ = [
    "PreparedCoo",
    "SDDMMPrepared",
    "PreparedCsrSpmv",
    "PreparedCsrSpmmOpt",
    "PreparedCsrSpmmOptAlg2",
    "SUPPORTED_INDEX_DTYPES",
    "SUPPORTED_VALUE_DTYPES",
    "benchmark_spmm_case",
    "benchmark_spmm_opt_alg2_case",
    "benchmark_spmm_opt_case",
    "benchmark_spmv_case",
    "comprehensive_spmm_test",
    "benchmark_sddmm_case",
    "flagsparse_sddmm_csr",
    "flagsparse_spmm_coo",
    "flagsparse_spmm_csr",
    "flagsparse_spmm_csr_opt",
    "flagsparse_spmm_csr_opt_alg2",
    "flagsparse_spmv_coo",
    "flagsparse_spmv_coo_tocsr",
    "flagsparse_spmv_csr",
    "prepare_sddmm_csr",
    "prepare_spmm_csr_opt",
    "prepare_spmm_csr_opt_alg2",
    "prepare_spmv_coo",
    "prepare_spmv_coo_tocsr",
    "prepare_spmv_csr",
]
