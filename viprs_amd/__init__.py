"""viprs_amd: MI355X-native executor for viprs's resample and separable
convolution path (shrink / reduce / thumbnail / resize / gauss_blur /
convsep), behind a C ABI (include/viprs_amd.h) with this Python mirror of the
reference's builder and operation API."""
from .api import (  # noqa: F401
    BICUBIC, BILINEAR, CATMULLROM, CancelToken, F32, FITBOX, FORCE_EXACT, HEIGHT, LANCZOS2, LANCZOS3, LBB,
    NEAREST, NOHALO, U8, U16, VSQBS, WIDTH, Batch, Context, DeviceBuffer, Op, Pipeline, Region,
    Server, VgpuError, thumbnail_shrink_on_load_factor,
)
from .api import E_SCRATCH_LIMIT, NOALLOC_SCRATCH_BYTES, F64, I16, I32, U32  # noqa: F401
from .api import (  # noqa: F401
    ENCODE_JPEG, ENCODE_PNG, JPEG_SUBSAMPLE_420, JPEG_SUBSAMPLE_AUTO, JPEG_SUBSAMPLE_OFF, jpeg_encode, png_decode,
    png_encode, process,
)
