"""MI355X (gfx950) pairwise-alignment engine: banded affine Smith-Waterman and
gap-affine WFA over 2-bit packed SequenceBatch buffers.

Host-side mirror of the reference's alignment interface (Computational-
Genomics-Mojo: genomics/core/sequence.mojo, genomics/cpu/align_cpu.mojo,
genomics/gpu/device.mojo, genomics/gpu/align_gpu.mojo) over the C ABI of
libcgm_align.so (include/cgm_align.h).  The package directory name is not a
Python identifier; load it with importlib (see tests/conftest.py).
"""
from . import _lib
from .align import (AlignConfig, AlignResult, MixedPlan, SwResult, WFAConfig, align_mixed_batch,
                    align_mixed_plain, default_config, default_tuning, default_wfa_config,
                    launch_banded_sw, launch_wfa_affine, set_default_tuning,
                    smith_waterman_banded, sw_banded_batch, wfa_affine, wfa_affine_batch,
                    wfa_affine_cpu, WFAResult, cigar_capacity, runs_to_cigar,
                    wfa_affine_cigar_batch, wfa_affine_cigar_cpu, needleman_wunsch_banded,
                    nw_banded_batch, SwStartResult, launch_banded_sw_start,
                    sw_banded_start_batch)
from .device import DeviceArray, GenomicsDevice, Timer
from .serve import GeneratorHandle, ServingEngine
from .sequence import (BASES_PER_WORD, Base, SequenceBatch, SequenceView, ascii_to_2bit,
                       concat_ascii, decode_ascii, encode_ascii_launch, encode_batch_gpu, get_base_from_word,
                       get_view, pack_codes)

__all__ = [
    "AlignConfig", "AlignResult", "MixedPlan", "align_mixed_batch", "align_mixed_plain",
    "SwResult", "WFAConfig", "default_config", "default_tuning", "default_wfa_config",
    "launch_banded_sw", "launch_wfa_affine", "set_default_tuning", "smith_waterman_banded",
    "sw_banded_batch", "wfa_affine", "wfa_affine_batch", "wfa_affine_cpu", "DeviceArray",
    "GenomicsDevice", "Timer", "GeneratorHandle", "ServingEngine", "BASES_PER_WORD", "Base",
    "SequenceBatch", "SequenceView", "ascii_to_2bit", "get_base_from_word", "get_view",
    "pack_codes", "concat_ascii", "decode_ascii", "WFAResult", "cigar_capacity",
    "runs_to_cigar", "wfa_affine_cigar_batch", "wfa_affine_cigar_cpu",
    "needleman_wunsch_banded", "nw_banded_batch", "encode_ascii_launch", "encode_batch_gpu",
    "SwStartResult", "launch_banded_sw_start", "sw_banded_start_batch",
]
