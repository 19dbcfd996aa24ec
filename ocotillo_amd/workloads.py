"""Seeded synthetic workloads (the benchmark configurations and test inputs).

All generators are deterministic (numpy PCG64, base seed 20261203) and
return C-ordered arrays in the solver's layout: rho, T as [ix, iy, z]
(Fortran (nz, nyloc, nxloc)).  The z grid is the reference's calc_grid
(source/disk.f90:42-62): nz points uniformly on [z0, z1].
"""
import numpy as np

SEED = 20261203
AU = 1.496e14              # H unit of the reference's disk samples (cm)
SIGMA_T = 2.54e14          # Gaussian T fall-off width of samples/NongreyAGN (cm)
T_FLOOR = 2000.0
# rho floor of the reference's non-grey samples (samples/NongreyAGN/input.in).
# A floor of 1e-24 g/cm^3 at 2000 K gives alpha ~ 1e-31 /cm, dz*alpha < 2^-52,
# and the reference's Thomas pivot becomes exactly zero (it stops).
RHO_FLOOR = 1e-15


def z_grid(nz, z0=-5e15, z1=5e15):
    dz = (z1 - z0) / (nz - 1)
    z = z0 + np.arange(nz) * dz
    return z, dz


def log_waves(nw, w0=100.0, w1=1e5):
    """nw wavelengths log-spaced on [w0, w1] Angstrom (100 A - 10 um by default)."""
    if nw == 1:
        return np.array([w0])
    return np.exp(np.linspace(np.log(w0), np.log(w1), nw))


def disk_columns(nx, ny, nz, seed=SEED + 1, rho_floor=RHO_FLOOR, z0=-5e15, z1=5e15, rng=None):
    """Non-grey disk columns: per column T_mid ~ U(2e4, 1e5) K with a Gaussian
    fall-off in z to a 2000 K floor, rho0 ~ logU(1e-11, 1e-8) g/cm^3,
    H ~ U(0.5, 2) * 1.496e14 cm, rho = max(rho0 exp(-z^2/2H^2), floor).
    Returns rho, T ([nx, ny, nz]) and dz."""
    rng = np.random.default_rng(seed) if rng is None else rng
    z, dz = z_grid(nz, z0, z1)
    ncol = nx * ny
    Tm = rng.uniform(2e4, 1e5, ncol)[:, None]
    rho0 = 10.0 ** rng.uniform(-11.0, -8.0, ncol)[:, None]
    H = rng.uniform(0.5, 2.0, ncol)[:, None] * AU
    T = np.maximum(Tm * np.exp(-z ** 2 / (2 * SIGMA_T ** 2)), T_FLOOR)
    rho = np.maximum(rho0 * np.exp(-z ** 2 / (2 * H ** 2)), rho_floor)
    return rho.reshape(nx, ny, nz), T.reshape(nx, ny, nz), dz


def config1():
    """Grey isothermal column (samples/GreyIsothermal): nz=10 on
    [-5e15, 5e15], rho = max(1e-9 exp(-z^2/2H^2), 1e-24), H = 1 AU, T = 2000 K
    (midTemp = isoTemp = 1 K is below the 2000 K floor), one wavelength,
    sigma_grey = 0.4."""
    z, dz = z_grid(10)
    rho = np.maximum(1e-9 * np.exp(-(0.5 * z * z) / (AU * AU)), 1e-24)
    T = np.full(10, 2000.0)
    return rho.reshape(1, 1, 10), T.reshape(1, 1, 10), dz, np.array([3000.0])


def config2(nx=64, ny=64, nz=256, nw=128):
    """Non-grey drop-in block: 64x64 columns, nz=256, nw=128 log-spaced
    100 A - 10 um, disk_columns profiles.  Returns rho, T, dz, waves."""
    rho, T, dz = disk_columns(nx, ny, nz, seed=SEED + 2)
    return rho, T, dz, log_waves(nw)


def config3(ncol=4096, seed=SEED + 3, nz_range=(32, 1024), nw_range=(16, 256), sigma_dex=0.5):
    """Batched uneven columns: nz ~ U{32..1024}, nw ~ U{16..256} (each column
    its own log-spaced 100 A - 10 um grid), disk profiles as config2 with a
    per-cell lognormal rho perturbation of sigma_dex dex, then the rho floor.
    Returns nz, nw (int arrays), waves (concatenated), dz (per column), and
    rho, T concatenated over columns."""
    rng = np.random.default_rng(seed)
    nz = rng.integers(nz_range[0], nz_range[1] + 1, ncol)
    nw = rng.integers(nw_range[0], nw_range[1] + 1, ncol)
    waves, dzs, rhos, Ts = [], [], [], []
    for c in range(ncol):
        rho, T, dz = disk_columns(1, 1, int(nz[c]), rng=rng)
        rho = np.maximum(rho.reshape(-1) * 10.0 ** rng.normal(0.0, sigma_dex, int(nz[c])), RHO_FLOOR)
        rhos.append(rho)
        Ts.append(T.reshape(-1))
        dzs.append(dz)
        waves.append(log_waves(int(nw[c])))
    return (nz.astype(np.int32), nw.astype(np.int32), np.concatenate(waves), np.array(dzs),
            np.concatenate(rhos), np.concatenate(Ts))


T_MID = 61341.0            # midplane temperature of samples/NongreyAGN/input.in:14 (K)


def gaussian_random_field(shape, corr_cells, rng):
    """Unit-variance Gaussian random field with Gaussian covariance
    exp(-r^2 / (2 corr^2)) (r in cells): white noise filtered in Fourier space
    by exp(-k^2 corr^2 / 4), periodic, then normalised to zero mean and unit
    variance."""
    noise = rng.standard_normal(shape)
    f = np.fft.rfftn(noise)
    del noise
    k2 = np.zeros(f.shape)
    for ax, n in enumerate(shape):
        k = 2.0 * np.pi * (np.fft.rfftfreq(n) if ax == len(shape) - 1 else np.fft.fftfreq(n))
        sh = [1] * len(shape)
        sh[ax] = k.size
        k2 = k2 + (k * k).reshape(sh)
    f *= np.exp(-k2 * (corr_cells * corr_cells) / 4.0)
    del k2
    g = np.fft.irfftn(f, s=shape, axes=tuple(range(len(shape))))
    g -= g.mean()
    g /= g.std()
    return g


def config4(nx=256, ny=256, nz=512, nw=128, seed=SEED + 4, sigma_dex=1.0, corr_cells=8.0,
            rho0=1e-9, H=AU, rho_floor=RHO_FLOOR):
    """Athena-like snapshot for the mixed-precision comparison: Gaussian
    background rho0 exp(-z^2/2H^2) times a lognormal 3-D Gaussian random field
    (sigma_dex dex, correlation length corr_cells cells) over the whole
    [ix, iy, z] box, floored at rho_floor; T = T_MID (rho/rho0)^0.3 clipped
    to [2000, 2e5] K.  z on [-5e15, 5e15], nw log-spaced 100 A - 10 um.
    Returns rho, T ([nx, ny, nz]), dz and waves."""
    rng = np.random.default_rng(seed)
    z, dz = z_grid(nz)
    g = gaussian_random_field((nx, ny, nz), corr_cells, rng)
    rho = rho0 * np.exp(-z ** 2 / (2 * H ** 2)) * 10.0 ** (sigma_dex * g)
    del g
    np.maximum(rho, rho_floor, out=rho)
    T = np.clip(T_MID * (rho / rho0) ** 0.3, T_FLOOR, 2e5)
    return rho, T, dz, log_waves(nw)


def config5_shapes(njobs, seed=SEED + 5, nxy=(16, 256), nz=(64, 2048), nw=(16, 512)):
    """Job shapes of the serving workload: nx, ny ~ U{16..256},
    nz ~ U{64..2048}, nw ~ U{16..512} (mean 5.16e9 points per job)."""
    rng = np.random.default_rng(seed)
    return [(int(rng.integers(nxy[0], nxy[1] + 1)), int(rng.integers(nxy[0], nxy[1] + 1)),
             int(rng.integers(nz[0], nz[1] + 1)), int(rng.integers(nw[0], nw[1] + 1))) for _ in range(njobs)]


def config5_job(nx, ny, nz, nw, seed, sigma_dex=0.5):
    """One serving job: disk profiles with a lognormal rho perturbation as in
    config3, nw log-spaced 100 A - 10 um.  Returns rho, T ([nx, ny, nz]), dz,
    waves."""
    rng = np.random.default_rng(seed)
    rho, T, dz = disk_columns(nx, ny, nz, rng=rng)
    rho *= 10.0 ** (sigma_dex * rng.standard_normal(rho.shape))
    np.maximum(rho, RHO_FLOOR, out=rho)
    return rho, T, dz, log_waves(nw)


def lognormal_pool(n, seed=SEED + 6, sigma_dex=0.5):
    """n seeded factors 10^(sigma_dex N(0, 1)): the per-cell rho perturbation
    of the serving jobs, drawn once and shared by slices."""
    rng = np.random.default_rng(seed)
    return 10.0 ** (sigma_dex * rng.standard_normal(n))


def serving_job_params(nx, ny, nz, seed, pool_size):
    """The seeded draws of one serving job (see serving_job_fill): per column
    T_mid ~ U(2e4, 1e5) K, rho0 ~ logU(1e-11, 1e-8), H ~ U(0.5, 2) AU, and
    the offset of its lognormal factors in the pool.  Returns (Tm, rho0, H,
    off, dz), Tm, rho0, H as [ncol, 1] arrays."""
    rng = np.random.default_rng(seed)
    _, dz = z_grid(nz)
    ncol = nx * ny
    Tm = rng.uniform(2e4, 1e5, ncol)[:, None]
    rho0 = 10.0 ** rng.uniform(-11.0, -8.0, ncol)[:, None]
    H = rng.uniform(0.5, 2.0, ncol)[:, None] * AU
    off = int(rng.integers(0, pool_size - ncol * nz + 1))
    return Tm, rho0, H, off, dz


def serving_job_fill(rho_out, T_out, c0, c1, nz, params, pool):
    """Columns c0 .. c1 - 1 of one serving job's inputs, written into the
    flat arrays rho_out, T_out (the whole job's, layout [ix, iy, z], e.g.
    pinned host memory): the disk profiles of disk_columns times a lognormal
    factor per cell, a contiguous slice of `pool` (lognormal_pool) at the
    job's offset, then the rho floor.  Disjoint column ranges may be filled
    concurrently."""
    Tm, rho0, H, off, _ = params
    z, _ = z_grid(nz)
    n0, n1 = c0 * nz, c1 * nz
    r = rho_out[n0:n1].reshape(c1 - c0, nz)
    np.multiply(-0.5 * z * z, 1.0 / (H[c0:c1] * H[c0:c1]), out=r)
    np.exp(r, out=r)
    r *= rho0[c0:c1]
    r *= pool[off + n0:off + n1].reshape(c1 - c0, nz)
    np.maximum(r, RHO_FLOOR, out=r)
    t = T_out[n0:n1].reshape(c1 - c0, nz)
    np.multiply(Tm[c0:c1], np.exp(-z ** 2 / (2 * SIGMA_T ** 2))[None, :], out=t)
    np.maximum(t, T_FLOOR, out=t)


def serving_job_into(rho_out, T_out, nx, ny, nz, seed, pool):
    """One serving job's inputs (serving_job_params, serving_job_fill over
    all its columns).  Returns dz."""
    params = serving_job_params(nx, ny, nz, seed, pool.size)
    serving_job_fill(rho_out, T_out, 0, nx * ny, nz, params, pool)
    return params[4]


def poisson_arrivals(njobs, rate, seed=SEED + 5):
    """Arrival times (s) of a Poisson stream of `rate` jobs per second."""
    rng = np.random.default_rng(seed + 1000)
    return np.cumsum(rng.exponential(1.0 / rate, njobs))
