"""`.svdag` input (EncodedSVDAG::load + decode, encoded_svdag.cpp:43-66, 200-281):
the device decoder against the oracle's DAG and encoders."""
import numpy as np
import pytest

import oracle as O
import svdag_amd as S
from test_gpu_parity import SCENES

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["terrain7", "caves6", "random6", "clustered7", "single", "full8"])
def test_svdag_decode_round_trip(ctx, name):
    """decode(encode(DAG)) is the DAG, level by level in file order; re-encoding
    gives the same bytes and the symmetry pass continues from it like the
    reference's svbuilder does for an .svdag input."""
    codes, L = SCENES[name]()
    print(f"[case] decode {name} L={L}", flush=True)
    o = O.OracleOctree.from_morton(codes, L)
    o.to_dag()
    o.init_child_levels()
    f = o.encode(0)
    g = S.Octree.from_svdag(ctx, f)
    for l in range(L):
        np.testing.assert_array_equal(g.level(S.VIEW_DAG, l), o.level(l), err_msg=f"level {l}")
    assert g.counts()[0] == o.n_nodes
    g.encode(S.FMT_SVDAG)
    assert g.encoded_bytes(S.FMT_SVDAG) == f
    if L >= 3:
        g.encode(S.FMT_ESVDAG)
        assert g.encoded_bytes(S.FMT_ESVDAG) == o.encode(2)
    o.to_sdag()
    g.to_sdag()
    g.encode(S.FMT_SSVDAG)
    if L >= 3:
        assert g.encoded_bytes(S.FMT_SSVDAG) == o.encode(2)
    g.encode(S.FMT_USSVDAG)
    assert g.encoded_bytes(S.FMT_USSVDAG) == o.encode(1)


def test_svdag_decode_from_device(ctx):
    import torch
    codes, L = SCENES["terrain7"]()
    g0 = S.Octree.from_morton(ctx, codes, L)
    g0.to_dag()
    g0.encode(S.FMT_SVDAG)
    f = g0.encoded_bytes(S.FMT_SVDAG)
    d = torch.frombuffer(bytearray(f), dtype=torch.uint8).to("cuda")
    g = S.Octree.from_svdag(ctx, (d.data_ptr(), d.numel()))
    for l in range(L):
        np.testing.assert_array_equal(g.level(S.VIEW_DAG, l), g0.level(S.VIEW_DAG, l))
    with pytest.raises(S.SvdError):
        g.to_dag()                      # already a DAG (svbuilder skips toDAG for .svdag input)


def test_svdag_decode_rejects_bad_files(ctx):
    codes, L = SCENES["terrain7"]()
    o = O.OracleOctree.from_morton(codes, L)
    o.to_dag()
    o.init_child_levels()
    f = bytearray(o.encode(0))
    with pytest.raises(S.SvdError):
        S.Octree.from_svdag(ctx, bytes(f[:30]))                 # header cut
    with pytest.raises(S.SvdError):
        S.Octree.from_svdag(ctx, bytes(f[:-4]))                 # word array cut
    bad = bytearray(f)
    bad[44 + 4:48 + 4] = (0xFFFFFF0).to_bytes(4, "little")      # root's first pointer outside the file
    with pytest.raises(S.SvdError):
        S.Octree.from_svdag(ctx, bytes(bad))
    bad = bytearray(f)
    bad[28:32] = (1).to_bytes(4, "little")                      # one level
    with pytest.raises(S.SvdError):
        S.Octree.from_svdag(ctx, bytes(bad))
    extra = bytearray(f)                                        # a trailing unreferenced word
    n = int.from_bytes(extra[40:44], "little")
    extra[40:44] = (n + 1).to_bytes(4, "little")
    extra += (1).to_bytes(4, "little")
    with pytest.raises(S.SvdError):
        S.Octree.from_svdag(ctx, bytes(extra))


def _svdag_file(levels, words):
    import struct
    head = struct.pack("<6ffIII", 0, 0, 0, 1, 1, 1, 1.0, levels, 4, len(words))
    return head + struct.pack("<I", len(words)) + struct.pack(f"<{len(words)}I", *words)


def test_svdag_decode_rejects_empty_nodes_below_root(ctx):
    """A node without children below the root never comes out of the encoder
    (cleanEmptyNodes and toDAG drop them); a file that holds one is rejected at
    decode instead of reaching toSDAG.  The hand-made file tiles its levels
    exactly, so the empty node is the only fault."""
    # root (slots 1, 0) -> level 1: X at word 3, Y at word 5, both -> the leaf 0xFF at word 7
    good = _svdag_file(3, [0x03, 5, 3, 0x01, 7, 0x01, 7, 0xFF])
    g = S.Octree.from_svdag(ctx, good)
    assert g.level_size(S.VIEW_DAG, 1) == 2
    g.free()
    # X without children at word 3, Y at word 4 -> the leaf at word 6
    bad = _svdag_file(3, [0x03, 4, 3, 0x00, 0x01, 6, 0xFF])
    with pytest.raises(S.SvdError, match="without children"):
        S.Octree.from_svdag(ctx, bad)
    leaf0 = _svdag_file(3, [0x01, 2, 0x01, 4, 0x00])
    with pytest.raises(S.SvdError, match="empty leaf"):
        S.Octree.from_svdag(ctx, leaf0)
