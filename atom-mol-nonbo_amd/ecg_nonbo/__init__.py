"""Host side of the MI355X ECG pair path (ctypes mirror of the reference's
matelem / matform / linalg / workproc interfaces)."""
from .system import EcgSystem, make_system, expand_yhy, transposition_matrices, parse_young_operator
from .inout import read_inout, write_inout, format_inout, InoutFile, BbopStep, HistoryEntry
from .swapfile import read_swap, write_swap
from .basis import seeded_basis, named_system, SYSTEMS, SEED_BASE, vech_index
