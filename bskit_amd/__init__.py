"""bskit_amd: MI355X-native FFT bispectrum estimator.

Drop-in for bskit's FFTBispectrum path (``import bskit_amd as bskit``): the
same class, module functions, result dict and output files, computed by
hand-written gfx950 HIP kernels (libbskhip.so) on HBM-resident data.
"""
from .bins import (closed_triangle_indices, generate_bin_edge_list,  # noqa: F401
                   generate_equilateral_triangle_bin_list, generate_isosceles_triangle_bin_list,
                   generate_squeezed_triangle_bin_list, generate_triangle_bin_list,
                   geometric_k_mean)
from .main import (FFTBispectrum, bk_binned, bk_FFT_full, bk_FFT_grid_info,  # noqa: F401
                   bk_FFT_unnormalized_value, compute_bk_FFT_value, compute_k_means_on_grid,
                   compute_Nbin, k_field, number_field, pk_FFT, power_spectrum, PRECISIONS)
from .mesh import ArrayMesh, MeshSource, ParticleMesh, cast_source  # noqa: F401
from .pipeline import GridPipeline  # noqa: F401
from .subbox import (field_subbox_pm, field_subbox_slice, measure_subboxes,  # noqa: F401
                     subbox_index_to_multiindex, subbox_multiindex_to_index)
from . import io  # noqa: F401

__version__ = '0.1.0'
