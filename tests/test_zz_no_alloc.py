"""GPU: the no-allocation energy deposit against eph_grid_insert_energy.

eph_grid_insert_energy_noalloc takes its points from device memory and its
scratch from one fixed device array per GPU; on the same points it must leave
the grid exactly as eph_grid_insert_energy does (same kernels, same order of
every per-cell sum).  Two calls in a row, the second with fewer points, show
that what one call leaves in the shared scratch does not leak into the next.
"""
import numpy as np
import pytest

import eph_amd as ea
from hipmem import DeviceArray, hip

pytestmark = pytest.mark.gpu


def _points(rng, n, L):
    # some points outside the box: the cell index folds them back periodically
    x = rng.uniform(-0.2 * L, 1.2 * L, (3, n))
    E = rng.standard_normal(n) * 1e-3
    return x[0], x[1], x[2], E


def _pair(nx, L):
    box = [0.0, L, 0.0, L, 0.0, L]
    a = ea.Grid.const(nx, nx, nx, box, 300.0, 3e-6, 1.0, 0.1)
    b = ea.Grid.const(nx, nx, nx, box, 300.0, 3e-6, 1.0, 0.1)
    for g in (a, b):
        g.set_dt(1e-3)
    return a, b


def test_noalloc_deposit_matches_allocating_entry_point():
    print("no-alloc deposit: 12^3 grid, 20000 then 3000 points", flush=True)
    rng = np.random.default_rng(11)
    L = 12 * 3.615
    a, b = _pair(12, L)
    try:
        for n in (20000, 3000):
            x, y, z, E = _points(rng, n, L)
            a.insert_energy(x, y, z, E)
            bufs = [DeviceArray(v) for v in (x, y, z, E)]
            try:
                b.insert_energy_noalloc(n, *(d.addr for d in bufs))
                dA, dB = a.get("dT"), b.get("dT")  # eph_grid_get orders after the queued work
            finally:
                for d in bufs:
                    d.close()
            assert np.array_equal(dA, dB), f"n={n}: max diff {np.abs(dA - dB).max():.3e}"
            assert np.count_nonzero(dB) > 0
        # the solve sees the same deposit
        a.solve()
        b.solve()
        assert np.array_equal(a.get("T"), b.get("T"))
    finally:
        a.close()
        b.close()


def test_noalloc_crowded_cells_and_limits():
    # 2^3 cells: first about 8750 points per cell (beyond the workgroup sort,
    # so the ordered scan over all points runs), then about 250 (workgroup
    # sort); then a size beyond the scratch is refused
    print("no-alloc deposit: 2^3 grid, crowded cells; limit check", flush=True)
    rng = np.random.default_rng(5)
    L = 2 * 3.615
    a, b = _pair(2, L)
    try:
        for n in (70000, 2000):
            x, y, z, E = _points(rng, n, L)
            a.insert_energy(x, y, z, E)
            with DeviceArray(x) as dx, DeviceArray(y) as dy, DeviceArray(z) as dz, DeviceArray(E) as dE:
                b.insert_energy_noalloc(n, dx.addr, dy.addr, dz.addr, dE.addr)
                dA, dB = a.get("dT"), b.get("dT")
            assert np.array_equal(dA, dB)
        cap = ea.lib().eph_noalloc_scratch_bytes()
        assert cap == 20 << 20
        too_many = cap // 16 + 1
        with DeviceArray(np.zeros(1)) as d:
            with pytest.raises(ea.EphError) as ei:
                b.insert_energy_noalloc(too_many, d.addr, d.addr, d.addr, d.addr)
        assert ei.value.code == -1 and "scratch" in str(ei.value)
        # nothing was queued by the refused call
        assert np.array_equal(a.get("dT"), b.get("dT"))
        assert hip().hipDeviceSynchronize() == 0
    finally:
        a.close()
        b.close()
