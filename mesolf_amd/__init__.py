"""MI355X (gfx950) phase-space Richardson-Lucy light-field deconvolution.

Native code: ``libmesolf.so`` (HIP kernels + C ABI, see ``include/mesolf.h``),
built in-tree by ``python -m mesolf_amd.build``.  This package is the thin
host-side mirror of the reference's MATLAB-facing interface.
"""
from .lib import (  # noqa: F401
    Batch, BudgetExceeded, Cancelled, Context, InvalidInput, MatFile, MatFileError, MesolfError, Patch, PSF, Server,
    Unsupported,
    compact_block_layout, exported_symbols, lfm_convolution_segments, lfm_convolution_segments_device,
    load_library,
    mex_lfm_convolution_vec, LFM_reconstruction_module, summary_image_generation_module,
    summary_image_pixelwise, summary_image_global, carpet_patch_info, crop_movie_device, img_rectify_mod, rectify_size,
    determine_patch_info_aug,
    recon_theory_phase_space_LFM_with_filter,
)

__all__ = [
    "Context", "PSF", "Patch", "Batch", "Server", "MesolfError", "InvalidInput", "Unsupported", "BudgetExceeded", "Cancelled",
    "MatFile", "MatFileError",
    "load_library", "exported_symbols", "compact_block_layout",
    "recon_theory_phase_space_LFM_with_filter", "mex_lfm_convolution_vec",
    "lfm_convolution_segments", "lfm_convolution_segments_device", "LFM_reconstruction_module", "summary_image_generation_module",
    "summary_image_pixelwise", "summary_image_global", "carpet_patch_info", "crop_movie_device", "img_rectify_mod", "rectify_size",
    "determine_patch_info_aug",
]
