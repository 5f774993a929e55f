"""GPU: the encoders' bounded look-back wait fails loudly.

The host forces the give-up with the test hook: a low wait bound (2000
sleeps) and one unit that never publishes its status.  Its successors wait,
reach the bound, count it as empty and set the launch's error bit; every
kernel still exits.  The fix-up then poisons offsets[nb], so each encoder's
caller sees LECO_FIX_EDEVICE: the C ABI's leco_fix_encode_result, and the
Python mirrors of the LeCo, FOR / Delta, op_max and batched encoders, which
raise.  With the hook reset, the same encodes equal the oracle again."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WITHHOLD = 5      # unit (ticket) that never publishes
SPINS = 2000      # wait bound while the hook is set


def say(msg):
    print(f"[gpu-test] {msg}", file=sys.stderr, flush=True)


@pytest.fixture
def hooked():
    from leco_fix import _lib
    L = _lib.lib()
    assert L.leco_fix_test_lookback_hook(SPINS, WITHHOLD) == 0
    try:
        yield L
    finally:
        assert L.leco_fix_test_lookback_hook(0, (1 << 64) - 1) == 0


def _expect_edevice(fn):
    from leco_fix._lib import LECO_FIX_EDEVICE, LecoError
    with pytest.raises(LecoError) as e:
        fn()
    assert e.value.code == LECO_FIX_EDEVICE


def test_lookback_give_up_poisons_every_encoder(hooked, oracle_mod):
    import torch
    from leco_fix import Column, datagen
    from leco_fix._lib import LECO_FIX_EDEVICE, LECO_FIX_OFFSETS_POISON
    from leco_fix.baselines import BaseColumn
    from leco_fix.batch import Batch
    from leco_fix.column import _stream_handle
    from leco_fix.opmax import OpMaxColumn
    L = hooked
    v64 = datagen.sorted_pwl_u64(200_000, seed=3)
    d64 = torch.from_numpy(v64.view(np.int64)).to("cuda")
    v32 = datagen.linear_noise_u32(200_000, seed=3)
    d32 = torch.from_numpy(v32.view(np.int32)).to("cuda")

    # the C ABI: the asynchronous encode returns OK, encode_result reports the poison
    for bs in (4096, 256):  # one block per unit (workgroup body), and wave-encoded units
        say(f"lookback give-up: C ABI encode_column_u64 bs={bs}")
        n = d64.numel()
        nb = int(L.leco_fix_block_count(n, bs))
        bound = int(L.leco_fix_stream_bound(n, bs, 64))
        st = torch.empty(bound, dtype=torch.uint8, device="cuda")
        offs = torch.empty(nb + 1, dtype=torch.int64, device="cuda")
        prm = torch.empty(nb * 32, dtype=torch.uint8, device="cuda")
        s = _stream_handle(None)
        assert L.leco_fix_encode_column_u64(d64.data_ptr(), n, bs, st.data_ptr(), bound, offs.data_ptr(),
                                            prm.data_ptr(), s) == 0
        size = ctypes.c_uint64(0)
        assert L.leco_fix_encode_result(offs.data_ptr(), nb, ctypes.byref(size), s) == LECO_FIX_EDEVICE
        assert int(offs[nb].item()) & LECO_FIX_OFFSETS_POISON == LECO_FIX_OFFSETS_POISON
        assert int(offs[0].item()) == 0  # the ticket is back to 0 for the next encode

    say("lookback give-up: Python mirrors")
    _expect_edevice(lambda: Column.encode(d64, 4096))
    _expect_edevice(lambda: Column.encode(d32, 512))
    _expect_edevice(lambda: BaseColumn.encode("for", d64, 1024))
    _expect_edevice(lambda: BaseColumn.encode("delta", d32, 1024))
    _expect_edevice(lambda: OpMaxColumn.encode(d32, 2000))

    def batch():
        b = Batch([d64, d32[:150_000]], [4096, 256])
        try:
            b.encode()
            torch.cuda.synchronize()
            b.compressed_bytes()
        finally:
            b.close()
    _expect_edevice(batch)

    # reset: the same encodes are byte-identical to the oracle again
    assert L.leco_fix_test_lookback_hook(0, (1 << 64) - 1) == 0
    for vals, dev, bs in ((v64, d64, 4096), (v32, d32, 512)):
        say(f"lookback hook reset: encode width={vals.dtype} bs={bs}")
        col = Column.encode(dev, bs)
        want, _ = oracle_mod.encode_column(vals, bs, threads=8)
        assert col.stream_bytes() == want.tobytes()
        size = ctypes.c_uint64(0)
        assert L.leco_fix_encode_result(col.offsets.data_ptr(), col.num_blocks, ctypes.byref(size),
                                        _stream_handle(None)) == 0
        assert size.value == col.nbytes
