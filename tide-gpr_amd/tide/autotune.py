"""Launch-shape autotuner for the TM2D step kernels.

The forward and adjoint kernels come in several families and take two
launch parameters:

* shots per wavefront: family 0 puts one shot on each wavefront, a
  workgroup of G wavefronts sharing a tile (its coefficient lines) across G
  shots and summing the FP64 gradient in LDS; the shot-loop family gives
  each wavefront one tile and walks every shot of the call (G is then tiles
  per workgroup);
* time steps per launch: 1, 2 (two-step kernels that recompute a halo of
  two and touch the wavefield state once per two steps, the interior and
  border tiles as one kernel or two) or K = 3, 4 (a halo of K, border
  tiles first in one kernel; host-staged snapshots run at 2);
* rows marched per wavefront (the tile height), and for two-step launches
  the height of the border tiles (0: the interior height, or 1 or 2 rows).

The best combination depends on the grid, the shot count and the pass (the
adjoint carries more state per row).  :func:`tune_tm2d` times every
candidate on a synthetic problem of the caller's size with HIP events,
installs the fastest per pass and caches it per size class, so later calls
with a similar problem reuse the choice.
"""

from __future__ import annotations

import time
from collections.abc import Sequence

import torch

from . import _native
from .maxwell.tm2d import prepare_tm2d
from .utils import C0

# (rows, group, shot loop, steps per launch, interior/border split,
#  border tile rows)
FWD_CANDIDATES: tuple[tuple[int, int, int, int, int, int], ...] = (
    (2, 4, 0, 1, 0, 0), (4, 2, 0, 1, 0, 0), (4, 4, 0, 1, 0, 0), (8, 4, 0, 1, 0, 0),
    (16, 4, 0, 1, 0, 0), (2, 4, 1, 1, 0, 0), (4, 2, 1, 1, 0, 0),
    (2, 2, 0, 2, 0, 0), (4, 1, 0, 2, 0, 0), (4, 2, 0, 2, 0, 0), (4, 4, 0, 2, 0, 0),
    (8, 2, 0, 2, 0, 0), (4, 1, 0, 2, 0, 2), (4, 2, 0, 2, 0, 2),
    (4, 1, 0, 2, 1, 0), (4, 2, 0, 2, 1, 0), (4, 2, 0, 3, 0, 0), (8, 2, 0, 3, 0, 0))
ADJ_CANDIDATES: tuple[tuple[int, int, int, int, int, int], ...] = (
    (2, 2, 0, 1, 0, 0), (2, 4, 0, 1, 0, 0), (2, 8, 0, 1, 0, 0), (4, 4, 0, 1, 0, 0),
    (8, 2, 0, 1, 0, 0), (2, 4, 1, 1, 0, 0), (4, 2, 1, 1, 0, 0),
    (4, 4, 0, 2, 1, 0), (4, 8, 0, 2, 1, 0), (4, 4, 0, 2, 1, 2), (4, 8, 0, 2, 1, 2),
    (2, 4, 0, 2, 0, 0), (4, 2, 0, 2, 0, 0), (4, 4, 0, 2, 0, 0), (4, 8, 0, 2, 0, 0),
    (4, 4, 0, 2, 0, 2), (4, 8, 0, 2, 0, 2), (4, 4, 0, 3, 0, 2), (4, 2, 0, 3, 0, 2))

_cache: dict[tuple, tuple[tuple[int, ...], tuple[int, ...]]] = {}

# A candidate replaces the installed shape only when it is faster by more
# than this fraction (the timings carry a few percent of run-to-run noise,
# and near-ties should not flip the choice from run to run).
KEEP_MARGIN = 0.03


def _pick(times: dict, installed: tuple) -> tuple:
    best = min(times, key=times.get)
    if installed in times and times[installed] <= (1.0 + KEEP_MARGIN) * times[best]:
        return installed
    return best


def _install(fwd: tuple[int, ...], adj: tuple[int, ...]) -> None:
    """Install one forward and one adjoint candidate."""
    _native.set_tm2d_tuning(fwd[0], fwd[1], adj[0], adj[1])
    _native.set_tm2d_shot_loop(fwd[2], adj[2])
    _native.set_tm2d_time_block(fwd[3], adj[3])
    _native.set_tm2d_tile_split(fwd[4], adj[4])
    _native.set_tm2d_border_rows(fwd[5], adj[5])


def current() -> tuple[tuple[int, ...], tuple[int, ...]]:
    """The installed (forward, adjoint) candidates."""
    t = _native.get_tm2d_tuning()
    lp = _native.get_tm2d_shot_loop()
    st = _native.get_tm2d_time_block()
    sp = _native.get_tm2d_tile_split()
    br = _native.get_tm2d_border_rows()
    return ((t[0], t[1], int(lp[0]), st[0], int(sp[0]), br[0]),
            (t[2], t[3], int(lp[1]), st[1], int(sp[1]), br[1]))


def _size_class(ny: int, nx: int, n_shots: int, dtype: torch.dtype) -> tuple:
    """Grids within a factor sqrt(2) in cells and shot counts in the same
    power-of-two bucket share a tuning."""
    cells = ny * nx
    return (int(round(2 * torch.log2(torch.tensor(float(cells))).item())),
            max(0, int(n_shots).bit_length() - 1), str(dtype))


def _problem(ny: int, nx: int, n_shots: int, nt: int, dtype: torch.dtype, device: torch.device):
    pml = 20 if min(ny, nx) > 80 else 5
    my, mx = max(ny - 2 * pml - 1, 4), max(nx - 2 * pml - 1, 4)
    g = torch.Generator().manual_seed(11)
    eps = (1 + 8 * torch.rand(my, mx, generator=g)).to(dtype=dtype, device=device)
    sig = (0.01 * torch.rand(my, mx, generator=g)).to(dtype=dtype, device=device)
    dx = 0.01
    dt = 0.9 * dx / (2**0.5 * C0)
    amp = torch.zeros(n_shots, 1, nt, dtype=dtype, device=device)
    amp[:, :, 1] = 1.0
    src = torch.stack([torch.tensor([[1, (3 + 7 * s) % mx]]) for s in range(n_shots)]).to(device)
    rec = torch.stack([torch.tensor([[1, x] for x in range(0, mx, max(1, mx // 32))])
                       for _ in range(n_shots)]).to(device)
    return prepare_tm2d(eps, sig, torch.ones_like(eps), dx, dt, amp, src, rec, pml_width=pml)


def _time(fn, reps: int = 2, setup=None) -> float:
    """Best of `reps` timings of fn() in milliseconds; `setup` (zeroing the
    state buffers) runs before each, outside the timed region."""
    best = float("inf")
    for _ in range(reps):
        if setup is not None:
            setup()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    return best


def tune_tm2d(
    ny: int,
    nx: int,
    n_shots: int | Sequence[int],
    *,
    dtype: torch.dtype = torch.float32,
    nt: int = 400,
    max_store_bytes: int = 24 << 30,
    device: torch.device | str = "cuda:0",
    fwd_candidates: Sequence[tuple[int, ...]] = FWD_CANDIDATES,
    adj_candidates: Sequence[tuple[int, ...]] = ADJ_CANDIDATES,
    use_cache: bool = True,
) -> dict:
    """Pick and install the fastest (rows, group, shot loop, steps, split,
    border rows) -- keeping the installed shape when it is within
    ``KEEP_MARGIN`` of the fastest -- for the forward-with-snapshots and adjoint kernels on a padded
    ``ny x nx`` grid with ``n_shots`` shots per call.  ``n_shots`` may
    also be the list of a caller's shot chunks (e.g. ``[7, 7, 6, 6, 6]``):
    each candidate is then timed on every distinct chunk size and scored by
    the total over the list, since the best shots-per-workgroup depends on
    the chunk.  ``nt`` steps per timed call (long enough that per-call set-up
    weighs as little as in a real run) are cut so that the snapshot buffers
    of all chunk sizes stay within ``max_store_bytes``.  Returns the choice (``forward`` and ``adjoint`` candidates)
    and the timings (microseconds per time step, summed over the chunks)."""
    device = torch.device(device)
    chunks = [int(n) for n in n_shots] if hasattr(n_shots, "__len__") else [int(n_shots)]
    if not chunks or min(chunks) < 1:
        raise ValueError("n_shots must be a positive count or a list of them")
    sizes = sorted(set(chunks))
    key = _size_class(ny, nx, max(chunks), dtype)
    if len(sizes) > 1:
        key = key + (tuple(sizes),)
    if use_cache and key in _cache:
        fwd, adj = _cache[key]
        _install(fwd, adj)
        return {"forward": fwd, "adjoint": adj, "cached": True}
    t_start = time.perf_counter()
    per_step = 2 * dtype.itemsize * ny * nx * sum(sizes)  # two snapshot planes per shot
    nt = max(16, min(int(nt), int(max_store_bytes // max(per_step, 1))) // 2 * 2)
    lib = _native.load()
    P = _native.ptr
    stream = _native.stream_handle(device)
    fwd = _native.get_function("maxwell_tm", "forward_with_storage", 2, dtype)
    bwd = getattr(lib, f"tide_tm2d_{_native.dtype_name(dtype)}_backward_f64acc")
    bpe = dtype.itemsize

    def runners(n: int):
        st = _problem(ny, nx, n, nt, dtype, device)
        S, pny, pnx = st.n_shots, st.ny, st.nx
        stores = [torch.empty(nt, S, pny, pnx, device=device, dtype=dtype) for _ in range(2)]
        r = torch.zeros(nt, S, st.n_rec, device=device, dtype=dtype)
        gr = torch.ones_like(r)
        gf = torch.zeros(nt, S, st.n_src, device=device, dtype=dtype)
        gca = torch.zeros(pny, pnx, device=device, dtype=torch.float64)
        gcb = torch.zeros_like(gca)

        fields = [torch.empty_like(t) for t in st.fields]
        lam = [torch.empty(S, pny, pnx, device=device, dtype=dtype) for _ in range(7)]

        def reset():
            for t, t0 in zip(fields, st.fields):
                t.copy_(t0)
            for t in lam:
                t.zero_()

        def run_fwd():
            _native.check(fwd(
                P(st.ca), P(st.cb), P(st.cq), P(st.f), *[P(t) for t in fields], P(r),
                P(stores[0]), None, None, P(stores[1]), None, None, *[P(p) for p in st.profiles],
                P(st.sources_i), P(st.receivers_i), st.rdy, st.rdx, st.dt, nt, S, pny, pnx,
                st.n_src, st.n_rec, 1, 0, 0, pny * pnx * bpe, True, True, False, False, False, 0,
                *st.pml, 0, device.index or 0, stream, stream))

        def run_bwd():
            _native.check(bwd(
                P(st.ca), P(st.cb), P(st.cq), P(gr), *[P(t) for t in lam], P(stores[0]),
                None, None, P(stores[1]), None, None, P(gf), P(gca), P(gcb),
                *[P(p) for p in st.profiles], P(st.sources_i), P(st.receivers_i), st.rdy,
                st.rdx, nt, S, pny, pnx, st.n_src, st.n_rec, 1, 0, 0, pny * pnx * bpe, False,
                False, False, nt, *st.pml, 2, device.index or 0, stream, stream))

        return run_fwd, run_bwd, reset

    runs = [(chunks.count(n), *runners(n)) for n in sizes]

    # The gradient-slice set-up and fold of a call is timed with it (it
    # belongs to the shape), but its weight is that of a call of nt steps;
    # nt should be long enough that the per-step cost dominates.
    def run_fwd():
        return sum(c * _time(f, setup=z) for c, f, _, z in runs)

    def run_bwd():
        return sum(c * _time(b, setup=z) for c, _, b, z in runs)

    fwd_before, adj_before = current()
    f_times, a_times = {}, {}
    try:
        for _, f, b, z in runs:
            z()
            f()
            z()
            b()
        torch.cuda.synchronize()
        for cand in fwd_candidates:
            _install(tuple(cand), adj_before)
            f_times[tuple(cand)] = run_fwd()
        best_f = _pick(f_times, fwd_before)
        for cand in adj_candidates:
            _install(best_f, tuple(cand))
            a_times[tuple(cand)] = run_bwd()
        best_a = _pick(a_times, adj_before)
    except BaseException:
        # a candidate that fails (unsupported shape, out of memory) must not
        # leave the process-wide launch shape at that candidate
        _install(fwd_before, adj_before)
        raise
    _install(best_f, best_a)
    _cache[key] = (best_f, best_a)

    def label(k):
        return (f"{k[0]}x{k[1]}{'L' if k[2] else ''}{f'/{k[3]}' if k[3] > 1 else ''}"
                f"{'s' if k[4] else ''}{f'b{k[5]}' if k[5] else ''}")

    return {
        "forward": best_f,
        "adjoint": best_a,
        "cached": False,
        "chunks": chunks,
        "nt": nt,
        "chosen_us_per_step": round((f_times[best_f] + a_times[best_a]) * 1e3 / nt, 2),
        "forward_us_per_step": {label(k): round(v * 1e3 / nt, 2) for k, v in f_times.items()},
        "adjoint_us_per_step": {label(k): round(v * 1e3 / nt, 2) for k, v in a_times.items()},
        "seconds": round(time.perf_counter() - t_start, 2),
    }


__all__ = ["tune_tm2d", "current", "FWD_CANDIDATES", "ADJ_CANDIDATES"]
