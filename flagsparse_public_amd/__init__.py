"""MI355X-native (gfx950) CSR SpMV / SpMM with FlagSparse's operator interface.

The public names resolve lazily from ``sparse_operations`` (as the reference's
``flagsparse/__init__.py:145-156`` does), so ``import flagsparse_public_amd``
does not load the HIP library until an operator is used.
"""

__version__ = "0.1.0"

_OPS_EXPORTS = {
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
}

__all__ = sorted(_OPS_EXPORTS)


def __getattr__(name):
    if name in _OPS_EXPORTS:
        from . import sparse_operations as _ops
        return getattr(_ops, name)
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")


def __dir__():
    return sorted(set(globals()) | _OPS_EXPORTS)
