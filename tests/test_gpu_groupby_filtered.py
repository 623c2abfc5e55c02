"""GPU: filtered group-by aggregation (qo_groupby_agg_filtered) -- the filter ->
group-by -> aggregate chain of host/host/testAggAfterGB.cpp:313-427, fused --
against the oracle's filter followed by its group-by, on the direct path, the
one-pass partitioned path and beyond 2^20 groups, with the filter on its own
column and on the group key itself (as the reference filters Record.y)."""
import numpy as np
import pytest

from query_on_opencl_fpga_amd.device import DeviceBuffer

pytestmark = pytest.mark.gpu


def run_filtered(qo, key, val, filt, lo, hi, G):
    n = key.shape[0]
    dk, dv = DeviceBuffer.from_numpy(key), DeviceBuffer.from_numpy(val)
    df = dk if filt is key else DeviceBuffer.from_numpy(filt)
    outs = [DeviceBuffer(max(G, 1) * s) for s in (4, 8, 8, 4, 4)]
    num = DeviceBuffer(8)
    try:
        qo._lib.call("qo_groupby_agg_filtered", dk.ptr, dv.ptr, df.ptr, n, lo, hi, G, *[b.ptr for b in outs],
                     num.ptr, None)
        g = int(num.to_numpy(np.uint64, 1)[0])
        return [b.to_numpy(d, g) for b, d in zip(outs, (np.int32, np.int64, np.int64, np.int32, np.int32))]
    finally:
        for b in outs + [num, dk, dv] + ([] if filt is key else [df]):
            b.free()


@pytest.mark.parametrize("n,G,z,lo,hi", [
    (1, 4, 0.0, 0, 100),
    (100_001, 1024, 0.5, 100, 399),         # direct path
    (2_000_000, 1024, 1.0, -(1 << 31), 49), # direct path, skewed
    (2_000_000, 1 << 20, 0.5, 250, 749),    # one-pass partitioned path
    (3_000_000, 1 << 21, 0.0, 0, 499),      # beyond 2^20 groups: one-pass, tables spill to global
    (500_000, 50_000, 0.8, 1000, 2000),     # nothing passes
])
def test_filtered_groupby_matches_oracle(qo, orc, n, G, z, lo, hi):
    t = orc.zipf_table(G, z)
    key = orc.gen_zipf_keys(n, t, 21, 1, 7).astype(np.int32)
    val = orc.gen_uniform_u32(n, 21, 2, 0, 1 << 16).astype(np.int32)
    filt = orc.gen_uniform_u32(n, 21, 3, 0, 1000).astype(np.int32)
    got = run_filtered(qo, key, val, filt, lo, hi, G)
    sel = (filt >= lo) & (filt <= hi)
    exp = orc.groupby_agg(key[sel], val[sel])
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)


def test_filter_on_the_group_key(qo, orc):
    """The reference chain: range_selection on .y, then group-by on .y."""
    n, G = 1_500_000, 5000
    key = orc.gen_uniform_u32(n, 22, 0, 0, G).astype(np.int32)
    val = orc.gen_uniform_u32(n, 22, 1, 0, 0).astype(np.int32)
    got = run_filtered(qo, key, val, key, 0, 50, G)
    sel = (key >= 0) & (key <= 50)
    exp = orc.groupby_agg(key[sel], val[sel])
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)
