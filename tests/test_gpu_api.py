"""C-ABI state after failed calls.

oco_set_wavelengths that cannot allocate the new grid's table (a context
budget too small for it) returns OCO_E_OOM and leaves the context without a
grid, so later solves and queries return OCO_E_STATE instead of reading a
freed table; a later grid that fits works again."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_failed_grid_leaves_no_grid():
    import ocotillo_amd as oco
    from ocotillo_amd.workloads import disk_columns, log_waves
    rho, T, dz = disk_columns(1, 1, 32)
    with oco.Context(0, hbm_budget=1 << 20) as c:
        c.set_physics(2e4, False, 0.4)
        c.set_wavelengths(log_waves(8))                     # 1.5 KB of table
        U, V, A, d = c.solve_block(rho, T, dz)
        assert d.status == 0
        with pytest.raises(oco.OcoError) as e:
            c.set_wavelengths(log_waves(100000))            # 18 MB of table
        assert e.value.status == oco.E_OOM
        for call in (lambda: c.solve_block(rho, T, dz), lambda: c.opacity(rho.reshape(-1), T.reshape(-1))):
            with pytest.raises(oco.OcoError) as e:
                call()
            assert e.value.status == oco.E_STATE
        c.set_wavelengths(log_waves(8))
        U2, V2, A2, d2 = c.solve_block(rho, T, dz)
        assert d2.status == 0 and np.array_equal(U2, U) and np.array_equal(V2, V)
