"""The captured Chebyshev filter (hipGraph replay) against the eager filter.

Both contexts run the same kernels with the same launch shapes, so every
replay must reproduce the eager result bit for bit.  Between replays the
program does its own eager work (H psi on another block, a Lanczos bound)
and synchronises, and every replay gets a new block and new filter bounds
(the coefficients are uploaded per call and read by the graph)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_filter_graph_replays_match_eager(si8, dtype):
    import torch
    from pwdft_develop_amd.ops import Context
    print(f"case: si8_32 filter graph replays {dtype}", flush=True)
    tdt = torch.complex128 if dtype == "f64" else torch.complex64
    cg = Context(si8)
    ce = Context(si8)
    ce.set_graphs(False)
    rng = np.random.default_rng(7)
    other = cg.to_device(si8.make_psi(si8.nb), tdt)
    v0 = cg.to_device(si8.lanczos_start())
    for it in range(6):
        X0 = si8.psi + 0.05 * it * (rng.standard_normal(si8.psi.shape) + 1j * rng.standard_normal(si8.psi.shape))
        X0[:, 0] = X0[:, 0].real
        a, b, aL = 0.3 + 0.05 * it, 12.0 + it, -1.0 - 0.1 * it
        Xg = cg.to_device(X0, tdt)
        Xe = ce.to_device(X0, tdt)
        cg.cheb_filter(Xg, 10, a, b, aL)   # call 0 eager, 1 captured, 2.. replayed
        ce.cheb_filter(Xe, 10, a, b, aL)
        # eager work of the program between replays
        cg.apply_h(other)
        cg.lanczos(6, v0)
        torch.cuda.synchronize()
        g, e = Xg.cpu().numpy(), Xe.cpu().numpy()
        assert np.array_equal(g, e), (it, np.abs(g - e).max())
    # a changed shape (degree) is captured anew and still matches
    Xg = cg.to_device(si8.psi, tdt)
    Xe = ce.to_device(si8.psi, tdt)
    for _ in range(3):
        cg.cheb_filter(Xg, 7, 0.4, 11.0, -0.5)
        ce.cheb_filter(Xe, 7, 0.4, 11.0, -0.5)
        torch.cuda.synchronize()
        assert np.array_equal(Xg.cpu().numpy(), Xe.cpu().numpy())
    cg.close()
    ce.close()


def test_filter_graph_dropped_on_new_projectors(si8):
    """Reallocating the projectors invalidates the captured graph: the next
    filters use the new operator (checked against a graphs-off context)."""
    import torch
    from pwdft_develop_amd.ops import Context
    print("case: si8_32 filter graph invalidation", flush=True)
    cg = Context(si8)
    ce = Context(si8)
    ce.set_graphs(False)
    for _ in range(3):   # eager, capture, replay
        Xg = cg.to_device(si8.psi)
        cg.cheb_filter(Xg, 8, 0.3, 12.0, -1.0)
    moved = si8.atoms + 0.1
    for ctx in (cg, ce):
        ctx.set_atoms(moved)
    for _ in range(3):
        Xg = cg.to_device(si8.psi)
        Xe = ce.to_device(si8.psi)
        cg.cheb_filter(Xg, 8, 0.3, 12.0, -1.0)
        ce.cheb_filter(Xe, 8, 0.3, 12.0, -1.0)
        torch.cuda.synchronize()
        assert np.array_equal(Xg.cpu().numpy(), Xe.cpu().numpy())
    cg.close()
    ce.close()


def test_lanczos_graph_replays_match_eager(si8):
    """The Lanczos steps replayed from a graph (after one eager call and one
    capture of the same step count) equal the eager steps bit for bit, with
    a different start vector buffer on every call, eager filter work in
    between, and a new step count captured anew."""
    import torch
    from pwdft_develop_amd.ops import Context
    print("case: si8_32 lanczos graph replays", flush=True)
    cg = Context(si8)
    ce = Context(si8)
    ce.set_graphs(False)
    rng = np.random.default_rng(11)
    X = cg.to_device(si8.psi)
    for k in (8, 8, 8, 8, 5, 5, 5, 8):
        v = si8.lanczos_start() + 0.1 * (rng.standard_normal(si8.npw) + 1j * rng.standard_normal(si8.npw))
        v[0] = v[0].real
        tg, bg = cg.lanczos(k, cg.to_device(v))
        te, be = ce.lanczos(k, ce.to_device(v))
        assert np.array_equal(tg, te) and bg == be, (k, np.abs(tg - te).max(), bg - be)
        cg.cheb_filter(X, 4, 0.3, 12.0, -1.0)  # eager or replayed filter work between calls
        torch.cuda.synchronize()
    cg.close()
    ce.close()
