"""The blocked driver's launch schedule, checked on the host (no GPU).

chol_plan_potrf returns the exact launch sequence blocked_potrf would
enqueue (diagonal blocks, panels, trailing updates, carried updates and
zero fill, with the side-stream fork / mark / join steps).  A block-level
model replays it and checks what a correct right-/left-looking blocked
Cholesky needs:

* every lower block (I, J) of 128 x 128 receives the update of each earlier
  column block K < J exactly once, and all of them before its own diagonal
  block or panel step reads it;
* an update only reads column blocks whose panel is final;
* the first touch of a block reads A (first-touch flag set), every later one
  reads L (flag clear);
* a carried update inside a diagonal-block launch never writes the block
  that launch factors, except in the in-kernel form, where the kernel
  applies it first;
* work on the side stream is not touched (read or written) by any launch the
  caller's stream enqueues between the fork before it and the join after
  it, and does not read what they write;
* L's upper triangle is zeroed exactly once: by the side-stream fill or by
  consecutive row ranges of the diagonal-block launches covering [0, n).

The schedules cover partial last blocks (n = 300, 320, 1100: 2*128+44,
2*128+64, 8*128+76), every outer block width the tuner can pick, both
forms, aligned and unaligned leading dimensions and other CU counts.
"""
import itertools

import numpy as np
import pytest

from gpumode_cholesky_amd import _lib, autotune

NB = 128
DIAG, PANEL, SYRK, ZERO, FORK, MARK, JOIN = 1, 2, 3, 4, 5, 6, 7


def plan(batch, n, ld, mode, cus=256):
    lib = _lib.load()
    import ctypes
    count = ctypes.c_int64(0)
    rc = lib.chol_plan_potrf(batch, n, ld, mode, cus, None, 0, ctypes.byref(count))
    assert rc == 0, _lib.last_error()
    rec = np.zeros((max(count.value, 1), 10), dtype=np.int64)
    rc = lib.chol_plan_potrf(batch, n, ld, mode, cus, rec.ctypes.data, count.value, ctypes.byref(count))
    assert rc == 0
    return rec[:count.value]


class Model:
    def __init__(self, n):
        self.n = n
        self.nb = (n + NB - 1) // NB
        nb = self.nb
        self.cnt = np.zeros((nb, nb, nb), np.int16)      # cnt[I, J, K]: updates by K applied to (I, J)
        self.written = np.zeros((nb, nb), bool)          # L holds the block (not A)
        self.diag_done = np.zeros(nb, bool)
        self.panel_done = np.zeros(nb, bool)
        self.lower = np.tril(np.ones((nb, nb), bool))
        self.zero_rows = []
        self.zero_side = 0
        # side-stream launches not yet joined: (reads, writes) block masks
        self.side_open = []
        self.side_marked = 0

    # ---- helpers
    def cols(self, c0, c1):
        assert c0 % NB == 0, c0
        assert c1 == self.n or c1 % NB == 0, c1
        return range(c0 // NB, (c1 + NB - 1) // NB)

    def kblocks(self, k0, kl):
        assert k0 % NB == 0 and kl % NB == 0 and kl > 0, (k0, kl)
        return range(k0 // NB, (k0 + kl) // NB)

    def check_first(self, blocks, first, what):
        w = [self.written[i, j] for (i, j) in blocks]
        assert all(w) or not any(w), f"{what}: blocks partly written, one first-touch flag"
        if blocks:
            assert bool(first) == (not w[0]), f"{what}: first-touch flag {first} but written={w[0]}"

    def update(self, c0, c1, k0, kl, first, what):
        """C -= P P^T over the lower blocks of columns [c0, c1), K = [k0, k0+kl)."""
        if c1 <= c0 or c0 >= self.n:
            return set(), set()
        cj = list(self.cols(c0, c1))
        ks = list(self.kblocks(k0, kl))
        tgt = [(i, j) for j in cj for i in range(j, self.nb)]
        for kb in ks:
            assert self.panel_done[kb] or (kb == self.nb - 1 and self.diag_done[kb]), \
                f"{what}: reads column block {kb} before its panel is final"
            assert kb < cj[0], f"{what}: K block {kb} not left of the updated columns"
        for j in cj:
            assert not self.diag_done[j], f"{what}: updates column block {j} after it was factored"
        self.check_first(tgt, first, what)
        for (i, j) in tgt:
            for kb in ks:
                self.cnt[i, j, kb] += 1
            self.written[i, j] = True
        reads = set(tgt) | {(i, kb) for kb in ks for i in range(cj[0], self.nb)}
        return reads, set(tgt)

    def touch(self, reads, writes, what):
        """A caller-stream launch against every unjoined side-stream launch."""
        for (sr, sw, swhat) in self.side_open:
            assert not (writes & (sr | sw)), f"{what} writes what {swhat} (side stream) touches"
            assert not (reads & sw), f"{what} reads what {swhat} (side stream) writes"

    # ---- records
    def run(self, rec):
        pending = []  # side launches enqueued, applied at the join (latest point)
        for r in rec:
            kind, stream = int(r[0]), int(r[1])
            if kind == FORK:
                continue
            if kind == MARK:
                self.side_marked = len(self.side_open)
                continue
            if kind == JOIN:
                # the caller now waits for every side launch up to the last mark
                self.side_open = self.side_open[self.side_marked:]
                self.side_marked = 0
                continue
            if kind == ZERO:
                assert stream == 1
                self.zero_side += 1
                continue
            if kind == SYRK:
                c0, c1, k0, kl, first = (int(x) for x in r[2:7])
                what = f"syrk cols [{c0},{c1}) K [{k0},{k0 + kl})"
                if stream == 1:
                    rd, wr = self.update(c0, c1, k0, kl, first, what)
                    self.side_open.append((rd, wr, what))
                else:
                    rd, wr = self.update(c0, c1, k0, kl, first, what)
                    self.touch(rd, wr, what)
                continue
            assert stream == 0, r
            if kind == DIAG:
                k0, first, cc0, cc1, ck0, cfl, zr0, zr1 = (int(x) for x in r[2:10])
                j = k0 // NB
                what = f"diag k0={k0}"
                rd, wr = set(), set()
                if cc0 >= 0:
                    pre = bool(cfl & 2)
                    if pre:
                        assert cc0 == k0, f"{what}: in-kernel update must start at its own block"
                    else:
                        assert cc0 >= k0 + NB, f"{what}: carried update writes the block being factored"
                    rd, wr = self.update(cc0, cc1, ck0, NB, cfl & 1, what + " carried")
                assert not self.diag_done[j]
                got = self.cnt[j, j, :]
                assert (got[:j] == 1).all() and (got[j:] == 0).all(), \
                    f"{what}: block ({j},{j}) has updates {got[:j + 1].tolist()}"
                self.check_first([(j, j)], first, what)
                self.diag_done[j] = True
                self.written[j, j] = True
                rd = rd | {(j, j)}
                wr = wr | {(j, j)}
                self.touch(rd, wr, what)
                if zr0 >= 0:
                    self.zero_rows.append((zr0, zr1))
            elif kind == PANEL:
                k0, first = int(r[2]), int(r[3])
                j = k0 // NB
                what = f"panel k0={k0}"
                assert self.diag_done[j] and not self.panel_done[j]
                blocks = [(i, j) for i in range(j + 1, self.nb)]
                for (i, jj) in blocks:
                    got = self.cnt[i, jj, :]
                    assert (got[:j] == 1).all() and (got[j:] == 0).all(), \
                        f"{what}: block ({i},{jj}) has updates {got[:j + 1].tolist()}"
                self.check_first(blocks, first, what)
                for b in blocks:
                    self.written[b] = True
                self.panel_done[j] = True
                self.touch(set(blocks) | {(j, j)}, set(blocks), what)
            else:
                raise AssertionError(f"unknown record {r}")

    def finish(self):
        assert not self.side_open, "side-stream work never joined"
        assert self.diag_done.all(), "a diagonal block was never factored"
        assert self.panel_done[:-1].all(), "a panel was never solved"
        for j in range(self.nb):
            for i in range(j, self.nb):
                got = self.cnt[i, j, :]
                assert (got[:j] == 1).all() and (got[j:] == 0).all(), (i, j, got.tolist())
        assert self.written[self.lower].all()
        if self.n > NB:
            if self.zero_side:
                assert self.zero_side == 1 and not self.zero_rows
            else:
                rows = sorted(self.zero_rows)
                assert rows and rows[0][0] == 0 and rows[-1][1] == self.n, rows
                for (a, b), (c, d) in zip(rows, rows[1:]):
                    assert b == c, f"zero-fill ranges {rows} leave a gap or overlap"


def check(batch, n, ld, mode, cus=256):
    rec = plan(batch, n, ld, mode, cus)
    if _lib.load().chol_route(batch, n, mode, None) != 3:
        # another route (the fused kernel by default for 128 < n <= 512 and
        # a batch of 256 and more): no blocked plan
        assert len(rec) == 0
        return rec
    m = Model(n)
    m.run(rec)
    m.finish()
    return rec


ORDERS = [65, 100, 128, 129, 200, 256, 257, 300, 320, 383, 384, 512, 640, 700, 1001, 1024, 1100, 2048]


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("n", ORDERS)
def test_schedule_default_table(n, mode):
    for batch in (1, 2, 3, 8, 16, 60, 200, 640):
        for ld in (n, (n + 3) // 4 * 4):
            check(batch, n, ld, mode)


@pytest.mark.parametrize("nbo", [128, 256, 384, 512, 640, 1024, 2048])
@pytest.mark.parametrize("mode", [0, 1])
def test_schedule_every_outer_width(nbo, mode):
    try:
        for n in (300, 320, 1100, 1001, 2048, 2176):
            autotune.set_table([[n, n, 1, 1 << 20, nbo]])
            for batch in (1, 3, 8, 60):
                check(batch, n, n, mode)
    finally:
        autotune.load_table()


@pytest.mark.parametrize("cus", [80, 304, 1024])
def test_schedule_other_cu_counts(cus):
    for n, batch, mode in itertools.product((300, 1100, 2048), (1, 3, 60, 200), (0, 1)):
        check(batch, n, n, mode, cus)


@pytest.mark.parametrize("batch,n", [(4, 1024), (60, 1024), (2, 2048), (8, 2048), (1, 4096), (2, 4096),
                                     (1, 8192), (1, 16384), (1, 32768), (16, 512), (640, 512), (1, 24576),
                                     (2, 24576), (1, 24700)])
@pytest.mark.parametrize("mode", [0, 1])
def test_schedule_bench_shapes(batch, n, mode):
    rec = check(batch, n, n, mode)
    assert len(rec) > 0


def test_schedule_model_catches_planted_mistakes():
    # The model itself must fail on plausible driver mistakes, planted in
    # a recorded schedule: a dropped trailing update, a doubled one, a bulk
    # update left unjoined, and a wrong first-touch flag.
    rec = plan(2, 1100, 1100, 1)
    syrk_rows = [i for i, r in enumerate(rec) if r[0] == SYRK]
    assert syrk_rows
    for mutate in ("drop", "double", "first"):
        bad = rec.copy()
        i = syrk_rows[len(syrk_rows) // 2]
        if mutate == "drop":
            bad = np.delete(bad, i, axis=0)
        elif mutate == "double":
            bad = np.insert(bad, i, bad[i], axis=0)
        else:
            bad[i, 6] ^= 1
        with pytest.raises(AssertionError):
            m = Model(1100)
            m.run(bad)
            m.finish()
    joins = [i for i, r in enumerate(rec) if r[0] == JOIN]
    if joins:
        with pytest.raises(AssertionError):
            m = Model(1100)
            m.run(np.delete(rec, joins[0], axis=0))
            m.finish()


def test_route_and_counters_host_only():
    import ctypes
    lib = _lib.load()
    nbo = ctypes.c_int64(-1)
    assert lib.chol_route(4096, 32, 0, ctypes.byref(nbo)) == 1 and nbo.value == 0
    assert lib.chol_route(3, 300, 0, ctypes.byref(nbo)) == 3 and nbo.value % 128 == 0
    assert lib.chol_route(0, 300, 0, None) == 0
    out = (ctypes.c_int64 * 8)()
    assert lib.chol_route_counters(out, 8) == 7
