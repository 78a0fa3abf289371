"""Inputs of short_char_hermite: a stratified box (chi, S from sweep_cases) with a
temperature, electron pressure and their gradients, beta from the temperature."""
import numpy as np

from sweep_cases import make_case

H_K = 6.62607015e-27 / 1.380649e-16


def hermite_inputs(case, seed=0, isothermal=False):
    rng = np.random.default_rng(seed)
    nz, ny, nx, nf = case["chi"][1:-1].shape
    z = np.arange(nz, dtype=np.float64)[:, None, None]
    if isothermal:
        T = np.full((nz, ny, nx), 5800.0)
    else:
        T = 5000.0 + 60.0 * z + 50.0 * rng.standard_normal((nz, ny, nx))
    pe = (10 ** (1.5 - 0.08 * z)) * np.exp(0.05 * rng.standard_normal((nz, ny, nx)))
    dzc = case["ds"][2]
    g = lambda a, ax, d: (np.roll(a, -1, ax) - np.roll(a, 1, ax)) / (2 * d)
    dpe = [g(pe, 2, case["ds"][0]), g(pe, 1, case["ds"][1]), np.gradient(pe, dzc, axis=0)]
    lt = np.log(T)
    dlnt = [g(lt, 2, case["ds"][0]), g(lt, 1, case["ds"][1]), np.gradient(lt, dzc, axis=0)]
    if isothermal:
        dlnt = [np.zeros_like(T)] * 3
    nu = np.linspace(5.9e14, 6.0e14, nf)
    hknu = (H_K * nu).astype(np.float32)
    beta = np.exp(-hknu[None, None, None, :].astype(np.float64) / T[..., None]).astype(np.float32)
    f32 = lambda a: np.ascontiguousarray(a, np.float32)
    return dict(beta=beta, pe4=f32(pe), temp=f32(T), dpe=[f32(a) for a in dpe], dlnt=[f32(a) for a in dlnt],
                hknu=hknu)
