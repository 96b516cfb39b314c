"""Pins of the oracle's boundary-row / column networks (the reference's
Mesh::NNmodel_Type01/03/05/07, IPPI-LBM/mesh.cpp:2767-2898, 3128-3260).

* The forward pass against torch.nn (Linear(15,128) -> ReLU -> 6 x
  (Linear(128,128) -> ReLU) -> Linear(128,1)) in FP64 and FP32: the oracle
  evaluates in FP32 with fused multiply-adds.
* Which node each type writes and which five neighbours it reads, in the
  reference's input order, with networks that pass one chosen input
  through to the output: the table below is the reference's source
  (node rows and columns at mesh.cpp:2786, 2852, 3146, 3213; inputs at
  mesh.cpp:2788-2804, 2854-2870, 3149-3165, 3216-3232).
* The guard: outputs below 0.501 leave the node's LES value
  (mesh.cpp:2821-2827, 393), and a row network decides where it meets a
  column network (it is called after it, mesh.cpp:363-378).
"""
import numpy as np
import pytest
import torch

NIN, NH = 15, 128

# type: (nodes as a function of (nx, ny), neighbour offsets in input order)
REF = {
    1: (lambda nx, ny: [(0, y) for y in range(1, ny - 1)], [(0, 1), (1, 1), (1, 0), (1, -1), (0, -1)]),
    3: (lambda nx, ny: [(x, ny - 2) for x in range(nx)], [(1, 0), (1, -1), (0, -1), (-1, -1), (-1, 0)]),
    5: (lambda nx, ny: [(nx - 1, y) for y in range(1, ny - 1)], [(0, 1), (-1, 1), (-1, 0), (-1, -1), (0, -1)]),
    7: (lambda nx, ny: [(x, 1) for x in range(nx)], [(-1, 0), (-1, 1), (0, 1), (1, 1), (1, 0)]),
}


def random_params(seed, out_bias=0.55, out_scale=0.02):
    rng = np.random.default_rng(seed)
    p = []

    def lin(nin, nout, scale=1.0):
        a = np.sqrt(6.0 / (nin + nout))
        p.append(rng.uniform(-a, a, (nout, nin)).ravel() * scale)
        p.append(rng.uniform(-0.1, 0.1, nout))
    lin(NIN, NH)
    for _ in range(6):
        lin(NH, NH)
    p.append(rng.uniform(-1, 1, NH) * out_scale)
    p.append([out_bias])
    return np.concatenate(p)


def torch_forward(params, x, dtype):
    layers, off = [], 0
    for nin, nout in [(NIN, NH)] + [(NH, NH)] * 6 + [(NH, 1)]:
        lin = torch.nn.Linear(nin, nout).to(dtype)
        with torch.no_grad():
            lin.weight.copy_(torch.tensor(params[off:off + nin * nout].reshape(nout, nin), dtype=dtype))
            off += nin * nout
            lin.bias.copy_(torch.tensor(params[off:off + nout], dtype=dtype))
            off += nout
        layers += [lin, torch.nn.ReLU()]
    net = torch.nn.Sequential(*layers[:-1])
    with torch.no_grad():
        return float(net(torch.tensor(np.asarray(x), dtype=dtype)[None])[0, 0])


def test_forward_against_torch(oracle):
    assert oracle.BNN_NPARAM == 101249
    p = random_params(3, out_bias=0.3, out_scale=1.0)
    rng = np.random.default_rng(5)
    for _ in range(20):
        x = np.concatenate([[rng.uniform(-0.1, 0.15), rng.uniform(-0.05, 0.05), rng.uniform(0.98, 1.02)]
                            for _ in range(5)])
        x32 = x.astype(np.float32)
        got = oracle.bnn_forward(p, x32)
        ref64 = torch_forward(p.astype(np.float32).astype(np.float64), x32.astype(np.float64), torch.float64)
        ref32 = torch_forward(p, x32, torch.float32)
        assert abs(got - ref64) <= 2e-6 * (1.0 + abs(ref64))
        assert abs(got - ref32) <= 2e-6 * (1.0 + abs(ref32))


def passthrough(j, bias):
    """out = relu(x_j) + bias: W0[0][j] = 1, hidden W[0][0] = 1, W7[0] = 1."""
    p = np.zeros(101249)
    p[0 * NIN + j] = 1.0
    off = NIN * NH + NH
    for _ in range(6):
        p[off] = 1.0  # W[0][0]
        off += NH * NH + NH
    p[off] = 1.0
    p[off + NH] = bias
    return p


def make(oracle, nx=12, ny=9, bc="inflow", les=True):
    s = oracle.Sim(nx, ny, "bgk", bc, les=les, tau0=0.52, cs=0.15, u_in=(0.06, 0.0), plate_x=4)
    s.init_fields(11, 1.0, (0.06, 0.0), 5e-3)
    return s


@pytest.mark.parametrize("net_type", [1, 3, 5, 7])
@pytest.mark.parametrize("comp", [0, 2])  # u.x and rho of each neighbour
def test_nodes_and_neighbours(oracle, net_type, comp):
    nx, ny = 12, 9
    nodes, offs = REF[net_type]
    for j, (dx, dy) in enumerate(offs):
        s = make(oracle)
        rho, ux, uy = s.fields()
        field = (ux, uy, rho)[comp]
        s.set_bnn(net_type, passthrough(3 * j + comp, 0.6))
        assert s.step(1) == 0
        tau = s.tau()
        for (x, y) in nodes(nx, ny):
            v = np.float32(field[y + dy, (x + dx) % nx])  # the ring wraps x
            exp = float(np.float32(max(v, np.float32(0.0))) + np.float32(0.6))
            assert tau[y, x] == exp, (net_type, j, x, y)
        s.close()


def test_guard_and_row_over_column(oracle):
    nx, ny = 12, 9
    base = make(oracle)
    base.step(1)
    t_les = base.tau()
    # rejected everywhere (rho - 0.9 + bias < 0.501): the LES tau stays
    s = make(oracle)
    s.set_bnn(3, passthrough(2, -0.6))
    s.step(1)
    assert np.array_equal(s.tau(), t_les)
    s.close()
    # a column network accepting at (0, 1) and (0, ny-2), the row networks rejecting there: LES wins
    s = make(oracle)
    s.set_bnn(1, passthrough(2, 0.6))
    s.set_bnn(7, passthrough(2, -0.6))
    s.set_bnn(3, passthrough(2, -0.6))
    s.step(1)
    t = s.tau()
    assert t[1, 0] == t_les[1, 0] and t[ny - 2, 0] == t_les[ny - 2, 0]
    assert np.all(t[2:ny - 2, 0] != t_les[2:ny - 2, 0])  # the column's own nodes take its value
    s.close()
    base.close()


def test_boundary_networks_change_the_flow(oracle):
    """Accepted network values enter the collision: fields differ from the
    run without networks after a few steps, mass stays finite."""
    a, b = make(oracle, bc="flatplate"), make(oracle, bc="flatplate")
    for t in (1, 3, 5, 7):
        b.set_bnn(t, random_params(10 + t))
    a.step(20)
    assert b.step(20) == 0
    assert not np.array_equal(a.fields()[1], b.fields()[1])
    assert np.isfinite(b.mass())
    a.close()
    b.close()
