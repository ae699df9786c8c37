"""cns_advance_async captured into a hipGraph and replayed.

The async step keeps dt on the device and makes no host call that a stream capture
forbids (the Courant accumulator is cleared by a kernel, not a memset), so a caller can
capture it once and replay it.  Each replay must be exactly one more step: the arrays and
dt after replay i are bit-identical to the eager cns_advance sequence after step i.  The
replays are checked one by one, with eager work and a synchronisation between them."""
import pytest

from cns_nospec_amd import ops
from cns_nospec_amd.layout import BoxLayout
from cns_nospec_amd.state import sinusoid_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [24, 40])
def test_async_step_graph_replays_equal_eager_steps(device, n):
    import torch
    lay = BoxLayout.cube(n)
    U0 = torch.from_numpy(sinusoid_state(lay, 90 + n, mach=0.2, kmax=3)).to(device)
    nsteps = 4
    print(f"graph replay n={n}", flush=True)

    arrs = ops.allocate_variables(lay, device)
    arrs[0].copy_(U0)
    st = ops.Stepper(lay)
    dts, states = [], []
    for _ in range(nsteps):
        dts.append(st.new_advance(*arrs))
        states.append([a.clone() for a in arrs])
    st.close()

    garrs = ops.allocate_variables(lay, device)
    dt_dev = torch.zeros(1, dtype=torch.float64, device=device)
    cour = torch.zeros(1, dtype=torch.float64, device=device)
    gst = ops.Stepper(lay)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):  # one uncaptured call first
        garrs[0].copy_(U0)
        gst.advance_async(*garrs, dt_dev, cour)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        gst.advance_async(*garrs, dt_dev, cour)
    for a in garrs[1:]:
        a.zero_()
    garrs[0].copy_(U0)
    torch.cuda.synchronize()
    for i in range(nsteps):
        graph.replay()
        torch.cuda.synchronize()
        assert dt_dev.item() == dts[i], f"replay {i}: dt"
        for name, a, b in zip(("U", "Unew", "Q", "D", "F"), garrs, states[i]):
            assert torch.equal(a, b), f"replay {i}: {name}"
        _ = garrs[0].sum().item()  # eager work between replays
    del graph
    gst.close()
