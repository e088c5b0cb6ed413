"""Shared builders for the parity tests: the same physical problem set up once through
the product (semhip) and once through the oracle, from the same inputs."""
import numpy as np

import semhip
from semhip import engine as E
from oracle import mesh as om
from oracle import shell as osh
from oracle import solid as oso
from oracle import common as oc


def shell_pair(nex=4, ney=3, order=5, lx=None, ly=None, perturb=0.0, seed=0, dv=(0, 0, 0, 0, 0),
               stiff_scale=None, rho_scale=None):
    """(oracle ShellModel, product inputs dict).  Default element sizes 12 x 10 mm keep the
    edge midpoints on the 1e-6 m grid that quad2spectral rounds them to, so the elements
    stay affine (other sizes make the reference's mesh slightly non-affine)."""
    lx = 0.012 * nex if lx is None else lx
    ly = 0.010 * ney if ly is None else ly
    q, p = om.rect_quad_mesh(lx, ly, nex, ney)
    if perturb:
        rng = np.random.default_rng(seed)
        inner = (p[:, 0] > 0) & (p[:, 0] < lx) & (p[:, 1] > 0) & (p[:, 1] < ly)
        p = p.copy()
        p[inner, :2] += perturb * rng.uniform(-1, 1, (inner.sum(), 2)) * min(lx / nex, ly / ney)
    nodes, coords = om.quad2spectral(q, p, order)
    Q = osh.lamina_q(130e9, 10e9, 5e9, 5e9, 3.5e9, 0.3)
    lam = osh.laminate_shell([0.25e-3] * 8, [0, 90, 0, 90, 90, 0, 90, 0], 1570, **Q)
    coef = np.array([[lam[k] for k in osh.SHELL_COEFS]])
    model = osh.ShellModel(nodes, coords, order, coef, stiff_scale=stiff_scale, rho_scale=rho_scale,
                           dv_field=dv)
    return model, dict(nodes=nodes, coords=coords, coef=coef, n_xy=order + 1)


def solid_pair(nex=2, ney=2, nlay=2, order=4, order_z=None, lx=None, ly=None, h=2e-3, perturb=0.0,
               seed=0, angles=(0, 90, 0, 90, 90, 0, 90, 0), stiff_scale=None, rho_scale=None, dv=(0, 0, 0)):
    order_z = order if order_z is None else order_z
    lx = 0.01 * nex if lx is None else lx
    ly = 0.01 * ney if ly is None else ly
    q, p = om.rect_quad_mesh(lx, ly, nex, ney)
    if perturb:
        rng = np.random.default_rng(seed)
        inner = (p[:, 0] > 0) & (p[:, 0] < lx) & (p[:, 1] > 0) & (p[:, 1] < ly)
        p = p.copy()
        p[inner, :2] += perturb * rng.uniform(-1, 1, (inner.sum(), 2)) * min(lx / nex, ly / ney)
    n2, c2 = om.quad2spectral(q, p, order)
    n3, c3 = om.extrude_mesh(n2, c2, order + 1, order + 1, order_z + 1, [h / nlay] * nlay, nlay)
    Q = oso.lamina_q3d(130e9, 10e9, 10e9, 5e9, 5e9, 3.5e9, 0.3, 0.3, 10 / 7 - 1)
    lam = oso.laminate_solid_smeared([0.25e-3] * len(angles), angles, 1570, **Q)
    coef = np.array([[lam[k] for k in oso.SOLID_COEFS]])
    model = oso.SolidModel(n3, c3, order + 1, order_z + 1, coef, stiff_scale=stiff_scale, rho_scale=rho_scale,
                           dv_field=dv)
    return model, dict(nodes=n3, coords=c3, coef=coef, n_xy=order + 1, n_z=order_z + 1)


def rel_l2(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
