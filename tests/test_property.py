"""Property tests (hypothesis): random tiny volumes encoded by the compiled
reference encoder; the oracle's decode of every brick at every LOD equals
the reference encoder's own majority-vote pyramid, and the producer's bytes
equal the reference bytes."""
import struct

import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings
from hypothesis import strategies as st

import fcs
from oracle import oracle as O

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="reference library not built here")


def file_order_ids(data):
    """Brick IDs in the order the .csdb lists them (fileIO.cpp:15-99)."""
    n = struct.unpack_from("<I", data, 8)[0]
    p = 16 + 8 + 64 + 64
    ids = []
    for _ in range(n):
        bid, P = struct.unpack_from("<II", data, p)
        p += 8 + 8 * P
        E = struct.unpack_from("<I", data, p)[0]
        p += 4 + E + 4
        ids.append(bid)
    return ids


@st.composite
def volumes(draw):
    """A cube grid of 1 or 2 bricks per axis (the format's brick IDs are
    Morton codes of a power-of-two cube) with ragged volume edges padded by
    the encoder."""
    b = draw(st.sampled_from([2, 4, 8]))
    nb_axis = draw(st.sampled_from([1, 2]))
    dims = [draw(st.integers((nb_axis - 1) * b + 1, nb_axis * b)) for _ in range(3)]
    k = draw(st.integers(1, 6))
    wide = draw(st.booleans())
    alphabet = [0] + [draw(st.integers(1, (1 << 63) if wide else (1 << 32) - 1)) for _ in range(k - 1)]
    seed = draw(st.integers(0, 2**32 - 1))
    rng = np.random.default_rng(seed)
    # blocky structure (constant sub-blocks give stop bits) mixed with noise
    block = draw(st.sampled_from([1, 2, 4]))
    shape = [-(-d // block) for d in dims[::-1]]
    coarse = rng.integers(0, k, shape)
    v = np.kron(coarse, np.ones((block,) * 3, np.int64))[:dims[2], :dims[1], :dims[0]]
    noise = rng.random(v.shape) < draw(st.sampled_from([0.0, 0.05, 0.3]))
    v = np.where(noise, rng.integers(0, k, v.shape), v)
    return np.array(alphabet, np.uint64)[v], b


@settings(max_examples=60, deadline=None, suppress_health_check=[HealthCheck.too_slow])
@given(volumes())
def test_oracle_decode_equals_reference_pyramid(tmp_path_factory, vb):
    vol, b = vb
    path = str(tmp_path_factory.mktemp("p") / "v.csbd")
    data, pyr = O.ref_encode(vol, b, path, with_pyramid=True)
    assert fcs.encode_volume(vol, b, header_brick_size_64=True) == data
    ds = O.Dataset(data, b)
    L = b.bit_length() - 1
    ids = file_order_ids(data)
    assert sorted(ids) == list(range(ds.num_bricks))
    for fi, bid in enumerate(ids):
        for t in range(L + 1):
            out, s = ds.decode(bid, t)
            assert s == 0
            assert np.array_equal(out, pyr[fi][O.pyramid_offset(t):O.pyramid_offset(t + 1)]), (bid, t)
