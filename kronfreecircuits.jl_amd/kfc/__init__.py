"""kfc: the Kronecker-free apply and eigensolver loops of KronFreeCircuits.jl on
the MI355X, over the C ABI of libkfc.so (include/kfc.h)."""
from ._lib import (Context, AbortError, ArgumentError, BreakdownError, ConvergenceError, DeviceError, DimensionMismatch,
                   UnsupportedError, context, lib, missing_symbols)
from .bases import ChargeBasis, FockBasis, PeriodicGrid, PhaseGrid, fd_weights, localop
from .localops import BandedOp, DenseOp, DiagonalOp, LocalOp, ShiftOp, dense, imagpart, realpart
from .operator import AxisTerm, KronFreeOperator
from .solvers import (LOBPCG, Auto, ChebyshevFilter, Dense, Lanczos, Solver, residuals, resolve, solve, solve_batch, solve_batch_subspace, solve_chebyshev_batch, solve_dense_batch,
                      solve_stats)
from .tree import (MaterializedNode, Node, NumericSpec, System, circuit_terms, constant_term, materialize, operator,
                   port_matrix, port_operator)
from .evolution import DriveHamiltonian, InteractionApply, KronFreeApply, evolve_rk4
from .resolvent import Resolvent, apply_resolvent, complement_solve, resolvent

__all__ = ["Context", "AbortError", "ArgumentError", "BreakdownError", "ConvergenceError", "DeviceError", "DimensionMismatch",
           "UnsupportedError", "context", "lib", "missing_symbols", "ChargeBasis", "FockBasis", "PeriodicGrid",
           "PhaseGrid", "fd_weights", "localop", "BandedOp", "DenseOp", "DiagonalOp", "LocalOp", "ShiftOp", "dense",
           "imagpart", "realpart", "AxisTerm", "KronFreeOperator", "LOBPCG", "Auto", "ChebyshevFilter", "Dense",
           "Lanczos", "Solver", "residuals", "resolve", "solve", "solve_batch", "solve_batch_subspace", "solve_chebyshev_batch", "solve_dense_batch", "solve_stats", "MaterializedNode", "Node",
           "NumericSpec", "System", "circuit_terms", "constant_term", "materialize", "operator", "port_matrix",
           "port_operator", "DriveHamiltonian", "InteractionApply", "KronFreeApply", "evolve_rk4",
           "Resolvent", "apply_resolvent", "complement_solve", "resolvent"]
