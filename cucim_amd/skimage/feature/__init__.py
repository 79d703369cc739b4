from .corner import hessian_matrix, hessian_matrix_eigvals  # noqa: F401
