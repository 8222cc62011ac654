"""MI355X-native LDPC belief-propagation decoder, a drop-in for FeCl's fec::Ldpc decode path."""
from .ldpc import (autotune_cache, autotune_cache_clear, autotune_cache_stats, Approximate, BatchGraph, archive_load, archive_save, CheckMatrix, DecoderAlgorithm, DecoderOptions, DvbS2, Exact, Gallager,
                   Ldpc, encode, Linear, MsgFormat, Permutation, PunctureOptions, Server, puncturing, scaling_table, structure)

__all__ = ["autotune_cache", "autotune_cache_clear", "autotune_cache_stats", "Approximate", "BatchGraph", "archive_load", "archive_save", "CheckMatrix", "DecoderAlgorithm", "DecoderOptions", "DvbS2", "Exact", "Gallager",
           "Ldpc", "Linear", "MsgFormat", "Permutation", "PunctureOptions", "Server", "puncturing", "scaling_table",
           "structure", "encode"]
