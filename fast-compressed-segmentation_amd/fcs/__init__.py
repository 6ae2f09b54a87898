"""Python host binding of libfcs (the MI355X brick decoder and DDA marcher).

Mirrors the reference's JavaScript entry points for this path
(web/app.js:9-12): ``load_dataset_from_array_buffer`` (datasetLoader.js:39),
``decode_all_rans_bricks`` (decode.js:17), ``request_bricks`` / ``decode_all``
(the decode dispatch of renderLoop.js:134-181, 268-298) and ``render``
(the rayTrace pass, renderLoop.js:284-296). Every call goes through the C ABI
in include/fcs/fcs.h; there is no Python or CPU fallback.
"""
from ._lib import (  # noqa: F401
    FcsError,
    lib,
    lib_path,
    OPS_HEAD_AUTO,
    OPS_HEAD_OFF,
    OPS_HEAD_ON,
    ST_BAD_OP,
    ST_BAD_REQUEST,
    ST_CACHE_FULL,
    ST_NSYM_MISMATCH,
    ST_PALETTE_RANGE,
    ST_RANS_STATE,
    ST_STREAM_OVERRUN,
)
from .api import (  # noqa: F401
    Camera,
    Context,
    Frame,
    Volume,
    load_dataset_from_array_buffer,
    decode_all_rans_bricks,
    device_count,
)
from .producer import Generator, encode_volume, encode_generated, brick_pyramid  # noqa: F401
from .mlp import MLP, pack_bits  # noqa: F401
from . import paths, prefetch, serve  # noqa: F401
