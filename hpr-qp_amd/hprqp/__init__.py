"""hprqp -- HPR-QP on AMD MI355X (gfx950).

Host API mirroring HPRQP.jl (src/HPRQP.jl:27-33 exports): build a model with
build_from_mps / build_from_QAbc / build_from_mat / build_from_ABST /
build_from_Ab_lambda, set HPRQP_parameters, call optimize().  The iteration
runs in libhprqp (hand-written HIP kernels behind a C ABI, include/hprqp.h).
"""
from .model import (
    AbstractQOperator,
    DiagonalPlusLowRank_Q_operator,
    LASSO_Q_operator,
    QAP_Q_operator,
    QP_info_cpu,
    build_from_ABST,
    build_from_Ab_lambda,
    build_from_mat,
    build_from_mps,
    build_from_QAbc,
    qp_formulation,
)
from .solver import DeviceModel, DeviceQOperator, HPRQP_parameters, HPRQP_results, load_autosave, optimize, to_gpu

__all__ = [
    "AbstractQOperator",
    "DiagonalPlusLowRank_Q_operator",
    "LASSO_Q_operator",
    "QAP_Q_operator",
    "QP_info_cpu",
    "build_from_ABST",
    "build_from_Ab_lambda",
    "build_from_mat",
    "build_from_mps",
    "build_from_QAbc",
    "qp_formulation",
    "DeviceModel",
    "DeviceQOperator",
    "HPRQP_parameters",
    "HPRQP_results",
    "load_autosave",
    "optimize",
    "to_gpu",
]
