"""Plain numpy statements of lattice definitions, used as independent checks.

They are written from the definitions (action, plaquette, Polyakov loop,
even/odd layout), not from the oracle's code: arrays are indexed [t, z, y, x]
so that the C-order flat index equals lex = x + Nx*(y + Ny*(z + Nz*t)).
"""
import numpy as np

AX = {0: 3, 1: 2, 2: 1, 3: 0}  # direction mu -> numpy axis of [t, z, y, x]


def coords(dims):
    Nx, Ny, Nz, Nt = dims
    t, z, y, x = np.meshgrid(np.arange(Nt), np.arange(Nz), np.arange(Ny), np.arange(Nx), indexing="ij")
    return x, y, z, t


def eo_index(dims):
    """For every site (C-order [t,z,y,x]) the EO offset eo + parity*V/2."""
    x, y, z, t = coords(dims)
    V = int(np.prod(dims))
    lex = x + dims[0] * (y + dims[1] * (z + dims[2] * t))
    par = (x + y + z + t) & 1
    return (lex >> 1) + par * (V // 2)


def to_lex(lat, dims):
    """EO flat array (4V) -> theta[mu] as arrays [t, z, y, x]."""
    V = int(np.prod(dims))
    idx = eo_index(dims)
    return [lat[mu * V + idx] for mu in range(4)]


def from_lex(th, dims):
    V = int(np.prod(dims))
    idx = eo_index(dims)
    out = np.zeros(4 * V)
    for mu in range(4):
        out[mu * V + idx.ravel()] = th[mu].ravel()
    return out


def shift(a, mu, n=1):
    """a(x + n*mu) as an array over x."""
    return np.roll(a, -n, axis=AX[mu])


def plaquette_angles(th):
    out = {}
    for mu in range(4):
        for nu in range(mu + 1, 4):
            out[(mu, nu)] = th[mu] + shift(th[nu], mu) - shift(th[mu], nu) - th[nu]
    return out


def action(th, xi=1.0):
    """sum_x sum_{mu<nu} w cos P, w = 1/xi space-space, xi space-time."""
    s = 0.0
    for (mu, nu), p in plaquette_angles(th).items():
        w = xi if (mu == 3 or nu == 3) else 1.0 / xi
        s += w * np.cos(p).sum()
    return s


def staple_from_action(th, xi, site, mu):
    """S such that the link-local action is Re(e^{i theta} S): read off from the
    total action at theta = 0, pi/2, pi."""
    vals = []
    for phi in (0.0, np.pi / 2, np.pi):
        t2 = [a.copy() for a in th]
        t2[mu][site] = phi
        vals.append(action(t2, xi))
    a0, a1, a2 = vals
    sre = (a0 - a2) / 2
    c = (a0 + a2) / 2
    sim = c - a1
    return complex(sre, sim)


def polyakov_field(th):
    """P(x) = exp(i sum_t theta_t(x, t)) over spatial sites [z, y, x]."""
    return np.exp(1j * th[3].sum(axis=0))


def correlator(P, rmax):
    """C(R) = (1/(3 SV)) sum_x sum_dir P(x) conj P(x + R e_dir)."""
    out = []
    for R in range(1, rmax + 1):
        s = 0j
        for ax in (2, 1, 0):  # x, y, z axes of [z, y, x]
            s += (P * np.conj(np.roll(P, -R, axis=ax))).sum()
        out.append(s / (3 * P.size))
    return np.array(out)
