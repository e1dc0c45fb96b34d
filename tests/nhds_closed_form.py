"""Closed-form hot-plasma susceptibility of a drifting bi-Maxwellian (Stix,
Waves in Plasmas, Eq. 10-57 with the plasma dispersion function), in the
normalisation of ALPS's NHDS module (ALPS_NHDS.f90:59-375).  Written with
scipy's Faddeeva function (wofz) and exponentially scaled Bessel I (ive);
used only to pin the numerical integration of the CPU restatement."""
import numpy as np
from scipy.special import ive, wofz


def Zfun(zeta, kpos=True):
    return 1j * np.sqrt(np.pi) * wofz(zeta) if kpos else -1j * np.sqrt(np.pi) * wofz(-zeta)


def chi_bimax(x, kperp, kz, m, q, n_s, beta_s, alpha, pdrift, nmax=80):
    """chi of one species in ALPS units (chi0 * omega^2 vA^2), kperp_norm = T.

    beta_s = n_s * beta * tau_s relates the NHDS species beta to the parameters of a
    generate_distribution bi-Maxwellian table (thermal momentum^2 = beta m tau)."""
    Om = q / m
    ell2 = m / (n_s * q * q)
    vth = np.sqrt(beta_s / (n_s * m))
    vd = pdrift / m
    z = 0.5 * (kperp * vth / Om) ** 2 * alpha
    Y = np.zeros((3, 3), complex)
    for n in range(-nmax, nmax + 1):
        res = x - kz * vd - n * Om
        zeta = res / (kz * vth)
        Z = Zfun(zeta, kz > 0)
        An = (alpha - 1) + (1 / (kz * vth)) * (alpha * res + n * Om) * Z
        Bn = ((alpha * (x - n * Om) - (kz * vd - n * Om)) / kz
              + ((x - n * Om) * (alpha * res + n * Om) / (kz * kz * vth)) * Z)
        B = ive(abs(n), z)
        dB = 0.5 * (ive(abs(n + 1), z) + ive(abs(n - 1), z))
        Y[0, 0] += n * n * B * An / z
        Y[0, 1] += -1j * n * (B - dB) * An
        Y[0, 2] += kperp * n * B * Bn / (Om * z)
        Y[1, 1] += (n * n * B / z + 2 * z * B - 2 * z * dB) * An
        Y[1, 2] += 1j * kperp * (B - dB) * Bn / Om
        Y[2, 2] += 2 * (x - n * Om) * B * Bn / (kz * vth * vth * alpha)
    chi = Y / ell2
    chi[2, 2] += 2 * x * vd / (ell2 * kz * vth * vth * alpha)
    chi[1, 0] = -chi[0, 1]
    chi[2, 0] = chi[0, 2]
    chi[2, 1] = -chi[1, 2]
    return chi
