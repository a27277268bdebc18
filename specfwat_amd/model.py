"""Seeded synthetic models sampled at GLL points.

* Gaussian random perturbation field eps (std 2 %, correlation length 5 km):
  Philox white noise on a regular grid, filtered so the covariance is
  proportional to exp(-r^2 / l^2), renormalised to zero mean and the target
  standard deviation.  vp, vs and rho are multiplied by (1 + eps).
* Sampling at GLL points with interp3 (trilinear), restated from
  interp3_0_cr (src/shared/utils.f90:1162-1232) in single precision, as
  model_external_values does (src/patch/model_external_values.f90:250-258).
* kappa/mu/rho_vp/rho_vs as get_model_fwat (src/mesh_database_api/get_model_fwat.f90:305-317).
* HTI -> 21 c_ij as AnisoStruct init_iso + hti2aniso + to_cijkl + glob2cart
  (src/patch/aniso.f90:55-162).
"""
import numpy as np

F32 = np.float32


def gaussian_random_field(shape, spacing, corr_len, std, seed):
    """Zero-mean field with covariance ~ exp(-r^2/corr_len^2) on a regular grid."""
    rng = np.random.Generator(np.random.Philox(seed))
    white = rng.standard_normal(shape)
    # filter g(r) ~ exp(-2 r^2 / l^2): g*g has covariance exp(-r^2 / l^2)
    pad = [int(np.ceil(2 * corr_len / spacing)) for _ in shape]
    padded = [n + 2 * p for n, p in zip(shape, pad)]
    W = np.zeros(padded)
    W[pad[0]:pad[0] + shape[0], pad[1]:pad[1] + shape[1], pad[2]:pad[2] + shape[2]] = white
    k = [np.fft.fftfreq(n, d=spacing) * 2 * np.pi for n in padded]
    kx, ky, kz = np.meshgrid(*k, indexing="ij")
    # Fourier transform of exp(-2 r^2 / l^2) ~ exp(-k^2 l^2 / 8)
    G = np.exp(-(kx**2 + ky**2 + kz**2) * corr_len**2 / 8.0)
    f = np.real(np.fft.ifftn(np.fft.fftn(W) * G))
    f = f[pad[0]:pad[0] + shape[0], pad[1]:pad[1] + shape[1], pad[2]:pad[2] + shape[2]]
    f -= f.mean()
    f *= std / f.std()
    return f


def locate_bissection(valx, x):
    """Vectorised locate_bissection (utils.f90:1490-1520); 1-based, -1 outside."""
    valx = np.asarray(valx, np.float64)
    x = np.asarray(x, np.float64)
    n = len(valx)
    jl = np.searchsorted(valx, x, side="right")  # number of nodes <= x
    ind = jl.astype(np.int64)
    ind = np.where(x == valx[0], 1, ind)
    ind = np.where(x == valx[n - 1], n - 1, ind)
    ind = np.where((x < valx[0]) | (x > valx[n - 1]), -1, ind)
    return ind


def interp3(valx, valy, valz, valin, x, y, z, chunk=1 << 20):
    """interp3_0_cr: float32 trilinear interpolation; valin[ix, iy, iz]."""
    x = np.asarray(x, F32); y = np.asarray(y, F32); z = np.asarray(z, F32)
    if x.size > chunk:
        out = np.empty(x.shape, F32)
        for c0 in range(0, x.size, chunk):
            sl = slice(c0, c0 + chunk)
            out.ravel()[sl] = _interp3(valx, valy, valz, valin, x.ravel()[sl], y.ravel()[sl], z.ravel()[sl])
        return out
    return _interp3(valx, valy, valz, valin, x, y, z)


def _interp3(valx, valy, valz, valin, x, y, z):
    valx = np.asarray(valx, F32); valy = np.asarray(valy, F32); valz = np.asarray(valz, F32)
    x = np.asarray(x, F32); y = np.asarray(y, F32); z = np.asarray(z, F32)
    nx, ny, nz = len(valx), len(valy), len(valz)
    m = 2
    kx = np.minimum(np.maximum(locate_bissection(valx.astype(np.float64), x.astype(np.float64)) - (m - 1) // 2, 1), nx + 1 - m) - 1
    ky = np.minimum(np.maximum(locate_bissection(valy.astype(np.float64), y.astype(np.float64)) - (m - 1) // 2, 1), ny + 1 - m) - 1
    kz = np.minimum(np.maximum(locate_bissection(valz.astype(np.float64), z.astype(np.float64)) - (m - 1) // 2, 1), nz + 1 - m) - 1
    xix1 = x - valx[kx]; yiy1 = y - valy[ky]; ziz1 = z - valz[kz]
    xix2 = x - valx[kx + 1]; yiy2 = y - valy[ky + 1]; ziz2 = z - valz[kz + 1]
    one = F32(1.0)
    x1x2 = one / (valx[kx] - valx[kx + 1]); y1y2 = one / (valy[ky] - valy[ky + 1]); z1z2 = one / (valz[kz] - valz[kz + 1])
    x2x1 = one / (valx[kx + 1] - valx[kx]); y2y1 = one / (valy[ky + 1] - valy[ky]); z2z1 = one / (valz[kz + 1] - valz[kz])
    V = np.asarray(valin, F32)
    f111 = xix2 * yiy2 * ziz2 * x1x2 * y1y2 * z1z2 * V[kx, ky, kz]
    f112 = xix2 * yiy2 * ziz1 * x1x2 * y1y2 * z2z1 * V[kx, ky, kz + 1]
    f121 = xix2 * yiy1 * ziz2 * x1x2 * y2y1 * z1z2 * V[kx, ky + 1, kz]
    f122 = xix2 * yiy1 * ziz1 * x1x2 * y2y1 * z2z1 * V[kx, ky + 1, kz + 1]
    f211 = xix1 * yiy2 * ziz2 * x2x1 * y1y2 * z1z2 * V[kx + 1, ky, kz]
    f212 = xix1 * yiy2 * ziz1 * x2x1 * y1y2 * z2z1 * V[kx + 1, ky, kz + 1]
    f221 = xix1 * yiy1 * ziz2 * x2x1 * y2y1 * z1z2 * V[kx + 1, ky + 1, kz]
    f222 = xix1 * yiy1 * ziz1 * x2x1 * y2y1 * z2z1 * V[kx + 1, ky + 1, kz + 1]
    return ((((((f111 + f112) + f121) + f122) + f211) + f212) + f221) + f222


def perturbed_model(mesh, vp0=6000.0, vs0=3500.0, rho0=2700.0, std=0.02, corr_len=5000.0,
                    grid=250.0, seed=1234, smooth=False):
    """rho, vp, vs (float32, [nspec,125]) of the benchmark recipe.  smooth=True
    returns the unperturbed background (the starting model)."""
    nloc = (mesh.nspec, 125)
    if smooth:
        return (np.full(nloc, rho0, F32), np.full(nloc, vp0, F32), np.full(nloc, vs0, F32))
    gx = np.arange(0.0, mesh.lx + 0.5 * grid, grid)
    gy = np.arange(0.0, mesh.ly + 0.5 * grid, grid)
    gz = np.arange(-mesh.lz, 0.5 * grid, grid)
    eps = gaussian_random_field((len(gx), len(gy), len(gz)), grid, corr_len, std, seed)
    fac = (1.0 + eps).astype(F32)
    fglob = interp3(gx, gy, gz, fac, mesh.xstore, mesh.ystore, mesh.zstore)
    f = fglob[mesh.ibool0]
    return (F32(rho0) * f, F32(vp0) * f, F32(vs0) * f)


def moduli(rho, vp, vs):
    """kappa, mu, rho_vp, rho_vs in single precision (get_model_fwat.f90:305-317)."""
    rho = np.asarray(rho, F32); vp = np.asarray(vp, F32); vs = np.asarray(vs, F32)
    four_thirds = F32(4.0 / 3.0)
    kappa = rho * (vp * vp - four_thirds * vs * vs)
    mu = rho * vs * vs
    return kappa, mu, rho * vp, rho * vs


CIJ_NAMES = ["c11", "c12", "c13", "c14", "c15", "c16", "c22", "c23", "c24", "c25", "c26",
             "c33", "c34", "c35", "c36", "c44", "c45", "c46", "c55", "c56", "c66"]


def hti_cij(vp, vs, rho, gcp, gsp):
    """21 c_ij (code frame 1 East, 2 North, 3 up) of an HTI medium, float32,
    last axis in CIJ_NAMES order (aniso.f90:55-162)."""
    vp = np.asarray(vp, F32); vs = np.asarray(vs, F32); rho = np.asarray(rho, F32)
    gcp = np.asarray(gcp, F32); gsp = np.asarray(gsp, F32)
    A = rho * vp * vp
    C = rho * vp * vp
    AL = rho * vs * vs
    AN = rho * vs * vs
    F = F32(1.0) * (A - F32(2.0) * AL)
    Gc = gcp * AL
    Gs = gsp * AL
    z = np.zeros_like(A)
    Bc = Bs = Hc = Hs = Ec = Es = C1p = C1sv = C1sh = S1p = S1sv = S1sh = C3 = S3 = z
    two = F32(2.0)
    d11 = A + Ec + Bc
    d12 = A - two * AN - Ec
    d13 = F + Hc
    d14 = S3 + two * S1sh + two * S1p
    d15 = two * C1p + C3
    d16 = -Bs / two - Es
    d22 = A + Ec - Bc
    d23 = F - Hc
    d24 = two * S1p - S3
    d25 = two * C1p - two * C1sh - C3
    d26 = -Bs / two + Es
    d33 = C
    d34 = two * (S1p - S1sv)
    d35 = two * (C1p - C1sv)
    d36 = -Hs
    d44 = AL - Gc
    d45 = -Gs
    d46 = C1sh - C3
    d55 = AL + Gc
    d56 = S3 - S1sh
    d66 = AN - Ec
    out = [d22, d12, d23, -d25, d24, -d26, d11, d13, -d15, d14, -d16, d33, -d35, d34, -d36,
           d55, -d45, d56, d44, -d46, d66]
    return np.stack(out, axis=-1).astype(F32)


def hti_model(mesh, rho, vp, vs, seed=77, max_strength=0.04, block=10000.0):
    """Random HTI per block: strength U[0, max_strength], fast azimuth U[0, pi)
    per `block` metres (horizontal), gcp = s cos 2phi, gsp = s sin 2phi."""
    rng = np.random.Generator(np.random.Philox(seed))
    nbx = int(np.ceil(mesh.lx / block))
    nby = int(np.ceil(mesh.ly / block))
    s = rng.uniform(0.0, max_strength, (nbx, nby))
    phi = rng.uniform(0.0, np.pi, (nbx, nby))
    xg = mesh.xyz[:, 0][mesh.ibool0]
    yg = mesh.xyz[:, 1][mesh.ibool0]
    bx = np.minimum((xg / block).astype(int), nbx - 1)
    by = np.minimum((yg / block).astype(int), nby - 1)
    gcp = (s * np.cos(2 * phi))[bx, by].astype(F32)
    gsp = (s * np.sin(2 * phi))[bx, by].astype(F32)
    return hti_cij(vp, vs, rho, gcp, gsp)
