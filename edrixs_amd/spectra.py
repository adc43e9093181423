"""Continued-fraction spectra from Krylov poles.

``get_spectra_from_poles`` follows edrixs/plot_spectrum.py:9-61,
``merge_pole_dicts`` edrixs/plot_spectrum.py:64-93 and ``boltz_dist`` /
``kelvin_to_beta`` edrixs/utils.py:34-73.  With ``device=True`` each pole set
is evaluated by the libedx continued-fraction kernel.
"""
import numpy as np

KB_EV = 8.6173303e-5


def kelvin_to_beta(k):
    return 1.0 / (KB_EV * k)


def boltz_dist(gs, T):
    g = np.asarray(gs, dtype=float)
    b = kelvin_to_beta(T)
    w = np.exp(-b * (g - g.min()))
    return w / w.sum()


def cf_host(alpha, beta, norm, e_gs, omega, gamma):
    """-1/pi Im[norm / (z - a_1 - b_1^2 / (z - a_2 - ...))], z = w + i g + E."""
    z = np.asarray(omega, dtype=float) + 1j * np.asarray(gamma, dtype=float) + e_gs
    t = np.zeros(z.shape, dtype=complex)
    for j in range(len(alpha) - 1, 0, -1):
        t = beta[j - 1] ** 2 / (z - alpha[j] - t)
    t = 1.0 / (z - alpha[0] - t)
    return -1.0 / np.pi * np.imag(t) * norm


def get_spectra_from_poles(poles_dict, omega_mesh, gamma_mesh, temperature, device=False):
    om = np.asarray(omega_mesh, dtype=float)
    ga = np.broadcast_to(np.asarray(gamma_mesh, dtype=float), om.shape)
    spectra = np.zeros(len(om))
    w = boltz_dist(poles_dict["eigval"], temperature)
    for i in range(len(poles_dict["eigval"])):
        n = poles_dict["npoles"][i]
        a = np.asarray(poles_dict["alpha"][i][:n], dtype=float)
        b = np.asarray(poles_dict["beta"][i][:n], dtype=float)
        e, nrm = poles_dict["eigval"][i], poles_dict["norm"][i]
        if device:
            from . import core
            core.cf_spectrum(a, b, nrm, e, om, ga, weight=w[i], out=spectra)
        else:
            spectra += cf_host(a, b, nrm, e, om, ga) * w[i]
    return spectra


def merge_pole_dicts(list_pole_dict):
    out = {"eigval": [], "npoles": [], "norm": [], "alpha": [], "beta": []}
    for d in list_pole_dict:
        for k in out:
            out[k].extend(d[k])
    return out


def relative_l2(s, ref):
    s, ref = np.asarray(s), np.asarray(ref)
    return float(np.linalg.norm(s - ref) / np.linalg.norm(ref))
