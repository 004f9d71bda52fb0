"""Differentiable beam residual and the implicit derivatives of the device
beam's Newton solves (the structural half of the reference's coupled
adjoint, coupled/gradients/coupled.py:195-353 and :540-629, which apply
jax.vjp to the beam residual of structure/beam.py:1749-1843).

The forward solves run on the device (``BeamStructure``'s Newton loops); this
module restates the residual they drive to zero as batched torch expressions
so that torch's autograd gives its exact derivatives:

* SE(3) maps of algebra/so3.py and algebra/se3.py with the reference's
  small-angle switches (the value of each branch where it is used; the
  exact branch is evaluated on a guarded argument elsewhere so that its
  gradient stays finite);
* element configuration d = log(H_A^-1 H_B) (structure/beam.py:1482-1500),
  strain, internal forces -P^T K eps (structure/beam.py:1329-1345), the
  inertial and gyroscopic forces -M_t v_dot - C_L v (structure/utils.py
  _integrate_m_l, _integrate_c_t; C_L needs dQ/dd along d_dot, taken here
  with torch.func.jvp), dead loads rotated into the node frames
  (structure/beam.py:1387-1395);
* ``StaticBeamSolve`` / ``DynamicBeamSolve``: autograd Functions whose
  forward is the device Newton solve and whose backward solves the
  transposed Jacobian of the residual at the converged state (implicit
  function theorem) and pulls the adjoint back through the residual's
  inputs (loads, stiffness scalings, previous-step state).

The implicit derivative is exact for a converged solve; run the forward with
tight structural tolerances when gradients are wanted.
"""

from __future__ import annotations

import numpy as np
import torch

F64 = torch.float64
SMALL = 1e-10  # utils/constants.py:13 (SMALL_ANG_THRESH)
COS_SMALL = float(np.cos(SMALL))
LOBATTO_MID = 0.5
LEGENDRE_S = (0.5 - 0.5 * np.sqrt(3.0 / 5.0), 0.5, 0.5 + 0.5 * np.sqrt(3.0 / 5.0))
LEGENDRE_W = (5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0)


# ---------------------------------------------------------------------------
# so(3) / se(3), algebra/so3.py and algebra/se3.py; vectors h = (u, omega)
# ---------------------------------------------------------------------------
def skew(v):
    z = torch.zeros_like(v[..., 0])
    return torch.stack((torch.stack((z, -v[..., 2], v[..., 1]), -1),
                        torch.stack((v[..., 2], z, -v[..., 0]), -1),
                        torch.stack((-v[..., 1], v[..., 0], z), -1)), -2)


def vee(m):
    return 0.5 * torch.stack((m[..., 2, 1] - m[..., 1, 2], m[..., 0, 2] - m[..., 2, 0], m[..., 1, 0] - m[..., 0, 1]),
                             -1)


def _n2(v):
    return (v * v).sum(-1)


def _eye(n, like):
    return torch.eye(n, dtype=like.dtype, device=like.device).expand(*like.shape[:-1], n, n)


def _guard(x, cond):
    """x where cond holds, 1 elsewhere (keeps the unused branch finite)."""
    return torch.where(cond, x, torch.ones_like(x))


def alpha_fn(b):
    n2 = _n2(b)
    big = n2 > SMALL * SMALL  # |b| > SMALL
    nb = torch.sqrt(_guard(n2, big))
    return torch.where(big, torch.sin(nb) / nb, 1.0 - n2 / 6.0)


def beta_fn(b):
    n2 = _n2(b)
    big = n2 > SMALL * SMALL
    n2g = _guard(n2, big)
    return torch.where(big, 2.0 * (1.0 - torch.cos(torch.sqrt(n2g))) / n2g, 0.5 - n2 / 24.0)


def t_so3(w):
    s = skew(w)
    n2 = _n2(w)
    big = n2 > SMALL
    n2g = _guard(n2, big)
    wg = torch.where(big[..., None], w, torch.ones_like(w))
    ss = s @ s
    exact = _eye(3, w) - 0.5 * beta_fn(wg)[..., None, None] * s + ((1.0 - alpha_fn(wg)) / n2g)[..., None, None] * ss
    return torch.where(big[..., None, None], exact, _eye(3, w) - s / 2.0 + ss / 6.0)


def t_inv_so3(w):
    s = skew(w)
    n2 = _n2(w)
    big = n2 > SMALL * SMALL  # |w| > SMALL
    n2g = _guard(n2, big)
    wg = torch.where(big[..., None], w, torch.ones_like(w))
    ss = s @ s
    exact = _eye(3, w) + 0.5 * s + ((1.0 - alpha_fn(wg) / beta_fn(wg)) / n2g)[..., None, None] * ss
    return torch.where(big[..., None, None], exact, _eye(3, w) + s / 2.0 + ss / 12.0)


def exp_so3(w):
    s = skew(w)
    n2 = _n2(w)
    big = n2 > SMALL
    n2g = _guard(n2, big)
    a = torch.sqrt(n2g)
    ss = s @ s
    exact = _eye(3, w) + (torch.sin(a) / a)[..., None, None] * s + ((1.0 - torch.cos(a)) / n2g)[..., None, None] * ss
    return torch.where(big[..., None, None], exact, _eye(3, w) + s + ss / 2.0)


def log_so3(r):
    """algebra/so3.py:log_so3, switch on cos(theta) < cos(SMALL).  The angle
    is taken as atan2(|vee(R - R^T)|, cos theta), the same angle as the
    reference's arccos(cos theta) but with a derivative that stays accurate
    for small rotations."""
    ct = 0.5 * (r[..., 0, 0] + r[..., 1, 1] + r[..., 2, 2] - 1.0)
    big = ct < COS_SMALL
    sv = 0.5 * vee(r - r.transpose(-1, -2))  # sin(theta) n
    s2 = _n2(sv)
    sg = torch.sqrt(_guard(s2, big & (s2 > 0.0)))
    th = torch.atan2(sg, ct)
    fac = torch.where(big & (s2 > 0.0), th / sg, torch.ones_like(th))
    exact = fac[..., None] * sv
    g = r - _eye(3, r[..., 0])
    small = vee(g - (g @ g) / 2.0)
    return torch.where(big[..., None], exact, small)


def hat6(h):
    z = torch.zeros_like(skew(h[..., :3]))
    su, sw = skew(h[..., :3]), skew(h[..., 3:])
    return torch.cat((torch.cat((sw, su), -1), torch.cat((z, sw), -1)), -2)


def check6(h):
    z = torch.zeros_like(skew(h[..., :3]))
    su, sw = skew(h[..., :3]), skew(h[..., 3:])
    return torch.cat((torch.cat((z, -su), -1), torch.cat((-su, -sw), -1)), -2)


def tilde4(h):
    sw = skew(h[..., 3:])
    top = torch.cat((sw, h[..., :3, None]), -1)
    bottom = torch.zeros_like(top[..., :1, :])
    return torch.cat((top, bottom), -2)


def t_u_omega_plus(h):
    """algebra/se3.py:t_u_omega_plus (Sonneville 2013 eq. A.12)."""
    a, b = h[..., :3], h[..., 3:]
    n2 = _n2(b)
    big = n2 > SMALL
    n2g = _guard(n2, big)
    bg = torch.where(big[..., None], b, torch.ones_like(b))
    sa, sb = skew(a), skew(b)
    al, be = alpha_fn(bg), beta_fn(bg)
    ba = (b * a).sum(-1)
    c = lambda x: x[..., None, None]  # noqa: E731
    exact = (-0.5 * c(be) * sa + c((1.0 - al) / n2g) * (sa @ sb + sb @ sa)
             + c(ba / n2g) * (c(be - al) * sb + c(0.5 * be - 3.0 * (1.0 - al) / n2g) * (sb @ sb)))
    return torch.where(c(big), exact, -0.5 * sa)


def _block(tl, tr, br):
    z = torch.zeros_like(tl)
    return torch.cat((torch.cat((tl, tr), -1), torch.cat((z, br), -1)), -2)


def t_se3(h):
    big = _n2(h[..., 3:]) > SMALL
    hg = torch.where(big[..., None], h, torch.ones_like(h))
    t = t_so3(hg[..., 3:])
    exact = _block(t, t_u_omega_plus(hg), t)
    a = hat6(h)
    return torch.where(big[..., None, None], exact, _eye(6, h) - a / 2.0 + (a @ a) / 6.0)


def t_inv_se3(h):
    big = _n2(h[..., 3:]) > SMALL
    hg = torch.where(big[..., None], h, torch.ones_like(h))
    ti = t_inv_so3(hg[..., 3:])
    exact = _block(ti, -ti @ t_u_omega_plus(hg) @ ti, ti)
    a = hat6(h)
    return torch.where(big[..., None, None], exact, _eye(6, h) + a / 2.0 + (a @ a) / 12.0)


def exp_se3(h):
    big = _n2(h[..., 3:]) > SMALL
    hg = torch.where(big[..., None], h, torch.ones_like(h))
    r = exp_so3(hg[..., 3:])
    p = (t_so3(hg[..., 3:]).transpose(-1, -2) @ hg[..., :3, None])
    top = torch.cat((r, p), -1)
    bottom = torch.zeros_like(top[..., :1, :])
    bottom = torch.cat((bottom[..., :3], torch.ones_like(bottom[..., 3:])), -1)
    exact = torch.cat((top, bottom), -2)
    x = tilde4(h)
    return torch.where(big[..., None, None], exact, _eye(4, h[..., :4]) + x + (x @ x) / 2.0)


def log_se3(g):
    w = log_so3(g[..., :3, :3])
    big = _n2(w) > SMALL
    wg = torch.where(big[..., None], w, torch.ones_like(w))
    exact = torch.cat(((t_inv_so3(wg).transpose(-1, -2) @ g[..., :3, 3, None])[..., 0], w), -1)
    e = g - _eye(4, g[..., 0])
    lg = e - (e @ e) / 2.0
    small = torch.cat((lg[..., :3, 3], vee(lg[..., :3, :3])), -1)
    return torch.where(big[..., None], exact, small)


def hg_inv(g):
    rt = g[..., :3, :3].transpose(-1, -2)
    p = -(rt @ g[..., :3, 3, None])
    top = torch.cat((rt, p), -1)
    return torch.cat((top, g[..., 3:, :]), -2)


def chi(r):
    z = torch.zeros_like(r)
    return torch.cat((torch.cat((r, z), -1), torch.cat((z, r), -1)), -2)


def p_mat(d, ad):
    """algebra/se3.py:p: [-T^-1(-d) Ad, T^-1(d) Ad] (6 x 12)."""
    tp = t_inv_se3(d)
    tn = tp - hat6(d)
    return torch.cat((-(tn @ ad), tp @ ad), -1)


def q_mat(s, d, ad):
    """algebra/se3.py:q with t_star = s T(s d) T^-1(d)."""
    ts = s * (t_se3(s * d) @ t_inv_se3(d))
    return torch.cat(((_eye(6, d) - ts) @ ad, ts @ ad), -1)


# ---------------------------------------------------------------------------
# Beam residual, structure/beam.py:1749-1843 with structure/utils.py
# ---------------------------------------------------------------------------
class BeamTables:
    """The element data of a beam as torch tensors: connectivity (n_elem, 2),
    element triads o0 (n_elem, 3, 3, columns x y z), lengths l0, reference
    configurations d0, stiffness and mass tables k_cs / m_cs (n, 6, 6) with
    the per-element indices into them."""

    def __init__(self, n_nodes, conn, o0, l0, d0, k_cs, m_cs, k_index, m_index, device):
        dev = torch.device(device)
        t = lambda x: torch.as_tensor(np.asarray(x, dtype=np.float64), device=dev)  # noqa: E731
        self.n_nodes = int(n_nodes)
        self.conn = torch.as_tensor(np.asarray(conn, dtype=np.int64).reshape(-1, 2), device=dev)
        self.n_elem = self.conn.shape[0]
        o0 = t(o0)
        self.base = torch.zeros((self.n_elem, 4, 4), dtype=F64, device=dev)
        self.base[:, :3, :3] = o0
        self.base[:, 3, 3] = 1.0
        self.ad = chi(o0.transpose(-1, -2))
        self.l0, self.d0 = t(l0), t(d0)
        self.k_cs, self.m_cs = t(k_cs), t(m_cs)
        self.k_index = torch.as_tensor(np.asarray(k_index, dtype=np.int64), device=dev)
        self.m_index = torch.as_tensor(np.asarray(m_index, dtype=np.int64), device=dev)
        idx = torch.cat((6 * self.conn[:, :1] + torch.arange(6, device=dev), 6 * self.conn[:, 1:] + torch.arange(6, device=dev)),
                        1)
        self.elem_dofs = idx  # (n_elem, 12)
        # quadrature constants on the device once (a per-call host-to-device
        # copy would also stop the backward passes from being captured)
        self.s_quad = torch.tensor((LOBATTO_MID,) + tuple(LEGENDRE_S), dtype=F64, device=dev)[:, None, None]
        self.w_quad = torch.tensor(LEGENDRE_W, dtype=F64, device=dev)[:, None, None, None]
        self.eye6 = torch.eye(6, dtype=F64, device=dev)

    @classmethod
    def from_beam(cls, beam):
        """From a device ``BeamStructure`` with its design variables set."""
        return cls(beam.n_nodes, beam._conn, beam.o0, beam.l0, beam.d0, beam.k_cs, beam.m_cs, beam.k_cs_index,
                   beam.m_cs_index, beam.hg0.device)

    def elem_vec(self, x):
        return x.reshape(-1)[self.elem_dofs]

    def assemble(self, fe):
        out = torch.zeros(6 * self.n_nodes, dtype=fe.dtype, device=fe.device)
        return out.index_add(0, self.elem_dofs.reshape(-1), fe.reshape(-1))


def element_d(tb: BeamTables, hg):
    """Element configurations d = log(H_A^-1 H_B) (structure/beam.py:1482-1500)."""
    ha = hg[tb.conn[:, 0]] @ tb.base
    hb = hg[tb.conn[:, 1]] @ tb.base
    return log_se3(hg_inv(ha) @ hb)


def dead_to_local(f, hg):
    r = hg[..., :3, :3]
    return torch.cat(((r.transpose(-1, -2) @ f[..., :3, None])[..., 0], (r.transpose(-1, -2) @ f[..., 3:, None])[..., 0]),
                     -1)


def beam_residual(tb: BeamTables, hg, k_scale=None, f_follower=None, f_dead=None, v=None, v_dot=None):
    """Full residual vector (6 n_nodes) f_int (+ f_iner + f_gyr) + follower +
    R^T dead (structure/beam.py:1749-1843); zero at a converged solve's
    solve DOFs.  ``k_scale`` multiplies each cross-section stiffness of the
    k_cs table (None: 1); v, v_dot (n_nodes, 6) select the dynamic form."""
    d = element_d(tb, hg)
    k = tb.k_cs if k_scale is None else tb.k_cs * k_scale[:, None, None]
    ke = k[tb.k_index]
    ad = tb.ad
    tinv = t_inv_se3(d)
    p = torch.cat((-((tinv - hat6(d)) @ ad), tinv @ ad), -1)  # p_mat with T^-1(d) shared
    eps = (d - tb.d0) / tb.l0[:, None]
    fe = -(p.transpose(-1, -2) @ (ke @ eps[..., None]))[..., 0]
    if v is not None:
        me = tb.m_cs[tb.m_index]
        ve = tb.elem_vec(v)
        vde = tb.elem_vec(v_dot)
        l = tb.l0
        mad = ad.transpose(-1, -2) @ me @ ad
        z = torch.zeros_like(mad)
        f0 = torch.cat((torch.cat((mad, z), -1), torch.cat((z, z), -1)), -2)
        f1 = torch.cat((torch.cat((z, z), -1), torch.cat((z, mad), -1)), -2)
        # Q at the Lobatto midpoint and the three Legendre points in one batch,
        # with its derivative along d_dot (the reference's q_dot JVP)
        s_all, eye = tb.s_quad, tb.eye6

        def qs(dd):
            ts = s_all[..., None] * (t_se3(s_all * dd[None]) @ t_inv_se3(dd)[None])
            return torch.cat(((eye - ts) @ ad, ts @ ad), -1)

        d_dot = (p @ ve[..., None])[..., 0]
        q, qd = torch.func.jvp(qs, (d,), (d_dot,))
        qm = q[0]
        m_t = l[:, None, None] * (0.5 * (2.0 / 6.0 * (f0 + f1) + 4.0 / 3.0 * (qm.transpose(-1, -2) @ me @ qm)))
        ql, qdl = q[1:], qd[1:]
        u = (ql @ ve[None, ..., None])[..., 0]
        terms = ql.transpose(-1, -2) @ (me @ qdl - hat6(u).transpose(-1, -2) @ me @ ql)
        c_l = 0.5 * l[:, None, None] * (tb.w_quad * terms).sum(0)
        fe = fe - (m_t @ vde[..., None])[..., 0] - (c_l @ ve[..., None])[..., 0]
    f = tb.assemble(fe)
    if f_follower is not None:
        f = f + f_follower.reshape(-1)
    if f_dead is not None:
        f = f + dead_to_local(f_dead, hg).reshape(-1)
    return f


# ---------------------------------------------------------------------------
# Implicit derivatives of the device Newton solves
# ---------------------------------------------------------------------------
def lie_gradient(hg, g_hg):
    """The gradient of a scalar with matrix gradient g_hg at hg, expressed in
    the right-trivialised increments xi of hg exp(xi) (n_nodes, 6)."""
    with torch.enable_grad():
        xi = torch.zeros(hg.shape[:-2] + (6,), dtype=F64, device=hg.device, requires_grad=True)
        (g,) = torch.autograd.grad((hg.detach() @ exp_se3(xi) * g_hg).sum(), (xi,))
    return g


def element_jacobians(beam, hg, k_elem, v=None, v_dot=None):
    """Exact derivative blocks of the element forces at the configuration hg
    (device kernel fj_beam_element_jacobians): (fe (E, 12), df/dd (E, 12, 6),
    df/dv_e (E, 12, 12) or None, M_t (E, 12, 12) or None, P(d) (E, 6, 12),
    eps (E, 6)); v, v_dot (n_nodes, 6) select the dynamic form."""
    import ctypes

    from . import _lib as L

    n_e = beam.n_elem
    dyn = v is not None
    hg = hg.contiguous()
    k_elem = k_elem.contiguous()
    vv = v.contiguous() if dyn else None
    vd = v_dot.contiguous() if dyn else None

    def e(*sh):
        return torch.empty(sh, dtype=F64, device=hg.device)

    fe, jd, p, eps = e(n_e, 12), e(n_e, 12, 6), e(n_e, 6, 12), e(n_e, 6)
    jv, mt = (e(n_e, 12, 12), e(n_e, 12, 12)) if dyn else (None, None)
    rc = L.lib().fj_beam_element_jacobians(ctypes.c_void_p(ctypes.addressof(beam._model)), int(dyn), L.ptr(hg),
                                           L.ptr(vv), L.ptr(vd), L.ptr(k_elem), L.ptr(fe), L.ptr(jd), L.ptr(jv),
                                           L.ptr(mt), L.ptr(p), L.ptr(eps), L.stream())
    L.check(rc, "fj_beam_element_jacobians")
    return fe, jd, jv, mt, p, eps


def _node_blocks_add(mat, blocks):
    """mat (6N x 6N) plus blockdiag(blocks (N, 6, 6))."""
    n = blocks.shape[0]
    mv = mat.view(n, 6, n, 6)
    ar = torch.arange(n, device=mat.device)
    mv[ar, :, ar, :] = mv[ar, :, ar, :] + blocks
    return mat


def _dead_stiffness(hg, f_dead):
    """d(R^T f)/d(right increment) per node: rows u and w, columns w:
    skew(R^T f_u), skew(R^T f_w) (R^T f changes by (R^T f) x w)."""
    rt = hg[:, :3, :3].transpose(-1, -2)
    lu = (rt @ f_dead[:, :3, None])[..., 0]
    lw = (rt @ f_dead[:, 3:, None])[..., 0]
    z = torch.zeros((hg.shape[0], 3, 3), dtype=F64, device=hg.device)
    return torch.cat((torch.cat((z, skew(lu)), -1), torch.cat((z, skew(lw)), -1)), -2)


class GraphedCall:
    """fn(*inputs) replayed from one captured HIP graph.  The backward passes
    of the implicit solves are thousands of small element-wise kernels on
    tensors of fixed shapes, launch-bound when issued one by one from the
    host; captured once per solver, a call is one copy of the inputs into the
    graph's buffers, one replay and a copy of the outputs.  fn must issue the
    same kernels whatever the values (no host reads of device data): the
    backward passes below qualify (their jacobian colouring and band-solve
    choice depend on shapes only and are settled in the eager warm-up)."""

    def __init__(self, fn, inputs):
        self.fn = fn
        self.sig = self.signature(inputs)
        self.static_in = [x.detach().clone() for x in inputs]
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(2):
                fn(*self.static_in)
        torch.cuda.current_stream().wait_stream(side)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.static_out = fn(*self.static_in)

    @staticmethod
    def signature(inputs):
        return tuple((tuple(x.shape), x.dtype, x.device) for x in inputs)

    def __call__(self, *inputs):
        for buf, x in zip(self.static_in, inputs):
            buf.copy_(x)
        self.graph.replay()
        return tuple(None if o is None else o.clone() for o in self.static_out)


class BeamSolver:
    """What the implicit Functions need from a device beam: the element tables,
    the solve DOFs and the two device Newton solves.  ``use_graphs``: replay
    the implicit backward passes from captured graphs (GPU tensors only)."""

    use_graphs = True
    device_pullback = True

    def __init__(self, beam, prescribed_dofs=(), time_integrator=None):
        self.beam = beam
        self.tables = BeamTables.from_beam(beam)
        pre = beam.make_prescribed_dofs_tuple(prescribed_dofs)
        self.prescribed = pre
        self.sd, self.n_s = beam._solve_dofs(pre)
        self.sd_long = self.sd.long()
        self.ti = time_integrator
        n = beam.n_nodes
        dev = beam.hg0.device
        self._bufs = (torch.empty((n, 4, 4), dtype=F64, device=dev), torch.empty((n, 6, 6), dtype=F64, device=dev))
        self.iters = []
        self._graphs = {}

    def run_backward(self, key, fn, inputs):
        """fn(*inputs), from a captured graph where the tensors are on the GPU."""
        if not (self.use_graphs and inputs[0].is_cuda):
            return fn(*inputs)
        gc = self._graphs.get(key)
        if gc is None or gc.sig != GraphedCall.signature(inputs):
            gc = GraphedCall(fn, inputs)
            self._graphs[key] = gc
        return gc(*inputs)

    def static(self, f_dead, k_scale):
        self.beam.set_stiffness_scaling(k_scale)
        out = self.beam.static_solve(self.prescribed, f_ext_aero=f_dead.contiguous())
        self.iters.append(out.iters[0])
        return out.hg

    def dynamic(self, hg_prev, phi, v, v_dot, f_dead, k_scale):
        b = self.beam
        b.set_stiffness_scaling(k_scale)
        phi, v, v_dot = phi.clone().contiguous(), v.clone().contiguous(), v_dot.clone().contiguous()
        self.iters.append(b._newton_dynamic(hg_prev.contiguous(), phi, v, v_dot, None, f_dead.contiguous(), self.sd,
                                            self.n_s, self.ti, self._bufs))
        return phi

    def reduce(self, x):
        return x.reshape(-1)[self.sd_long]

    def solve_selector(self, device):
        """1 at the solve DOFs, 0 elsewhere (6 n_nodes)."""
        if getattr(self, "_sel", None) is None:
            sel = torch.zeros(6 * self.tables.n_nodes, dtype=F64)
            sel[self.sd_long.cpu()] = 1.0
            self._sel = sel.to(device)
        return self._sel

    def assemble_blocks(self, blocks):
        """Dense 6N x 6N matrix from element blocks (E, 12, 12) over the
        element DOFs (at most two contributions per entry: exact in any
        order)."""
        if getattr(self, "_asm", None) is None:
            idx = self.tables.elem_dofs
            n6 = 6 * self.tables.n_nodes
            self._asm = ((idx[:, :, None] * n6 + idx[:, None, :]).reshape(-1), n6)
        flat, n6 = self._asm
        out = torch.zeros(n6 * n6, dtype=F64, device=blocks.device)
        return out.index_add_(0, flat, blocks.reshape(-1)).view(n6, n6)

    def k_gradient(self, mu_full, p, eps):
        """sum over the elements of table entry s of mu_e . P^T K_s eps: the
        stiffness-scaling part of -mu^T dR/dk (the internal force is
        -P^T (k_s K_s) eps)."""
        tb = self.tables
        mu_e = mu_full[tb.elem_dofs]
        kb = tb.k_cs[tb.k_index]
        f = (p.transpose(-1, -2) @ (kb @ eps[..., None]))[..., 0]
        contrib = (mu_e * f).sum(-1)
        return torch.zeros(tb.k_cs.shape[0], dtype=F64, device=p.device).index_add_(0, tb.k_index, contrib)

    def expand(self, xs, like):
        out = torch.zeros(like.numel(), dtype=F64, device=like.device)
        return out.index_put((self.sd_long,), xs).reshape(like.shape)

    def _node_coloring(self):
        """Colours of the nodes such that no two nodes of one colour share an
        element neighbour (distance-2 colouring of the element graph): a
        residual row of node i depends only on the states of i and its
        element neighbours, so the columns of one colour never meet in a row."""
        n = self.tables.n_nodes
        nb = [{i} for i in range(n)]
        for a, b in self.tables.conn.tolist():
            nb[a].add(b)
            nb[b].add(a)
        colour = [-1] * n
        for i in range(n):
            used = {colour[k] for j in nb[i] for k in nb[j] if colour[k] >= 0}
            c = 0
            while c in used:
                c += 1
            colour[i] = c
        near = torch.zeros((n, n), dtype=torch.bool)
        for i in range(n):
            near[i, list(nb[i])] = True
        return torch.as_tensor(colour), near

    def jacobian(self, res, x):
        """d res / d x over the solve DOFs (n_s x n_s), exact, from one
        forward-mode product per (colour, DOF component) instead of one per
        column."""
        if getattr(self, "_colour", None) is None:
            colour, near = self._node_coloring()
            sd = self.sd_long.cpu()
            node, comp = sd // 6, sd % 6
            seed = colour[node] * 6 + comp
            n_col = int(seed.max()) + 1 if len(sd) else 0
            dev = x.device
            v = torch.zeros((self.n_s, n_col), dtype=F64)
            v[torch.arange(self.n_s), seed] = 1.0
            self._colour = (seed.to(dev), v.to(dev), near[node][:, node].to(dev))
        seed, v, mask = self._colour
        jv = torch.func.vmap(lambda t: torch.func.jvp(res, (x,), (t,))[1], in_dims=1, out_dims=1)(v)
        return torch.where(mask, jv[:, seed], torch.zeros((), dtype=F64, device=x.device))

    def solve_transposed(self, jac, g):
        from . import _kernels as K

        bw = getattr(self.beam, "_bandwidth", None) if hasattr(self, "beam") else None
        jt = jac.transpose(0, 1).contiguous()
        if jac.is_cuda and bw is not None and K.band_fits(self.n_s, bw, bw):
            return K.band_solve(jt, g.contiguous(), bw, bw)
        return torch.linalg.solve(jt, g)


class StaticBeamSolve(torch.autograd.Function):
    """hg = the static Newton solution under global dead loads f_dead
    (n_nodes, 6) with stiffness scalings k_scale (structure/beam.py:1878-2092
    on the device).  Backward: J^T mu = dL/dxi at the solution, J = dR/dxi of
    R(hg exp(xi)) over the solve DOFs, then -mu^T dR/d(f_dead, k_scale)
    (coupled/gradients/coupled.py:195-353)."""

    @staticmethod
    def forward(ctx, f_dead, k_scale, solver):
        hg = solver.static(f_dead.detach(), k_scale.detach())
        ctx.solver = solver
        ctx.save_for_backward(hg, f_dead, k_scale)
        return hg

    @staticmethod
    def backward(ctx, g_hg):
        hg, f_dead, k_scale = ctx.saved_tensors
        sv = ctx.solver
        inputs = (hg.detach(), f_dead.detach(), k_scale.detach(), g_hg.contiguous())
        fn = StaticBeamSolve.pullback_device if (sv.device_pullback and hg.is_cuda) else StaticBeamSolve.pullback
        gf, gk = sv.run_backward("static", lambda *x: fn(sv, *x), inputs)
        return gf, gk, None

    @staticmethod
    def pullback_device(sv, hg, fd, ks, g_hg):
        """pullback from the exact element derivative blocks of the device
        kernel: J = dR/dxi is the assembled dfe/dd P plus the dead-load
        blocks (xi = 0 at the solution, so no tangent factor)."""
        tb = sv.tables
        g = sv.reduce(lie_gradient(hg, g_hg))
        k_elem = (tb.k_cs * ks[:, None, None])[tb.k_index]
        _, jd, _, _, p, eps = element_jacobians(sv.beam, hg, k_elem)
        jz = _node_blocks_add(sv.assemble_blocks(jd @ p), _dead_stiffness(hg, fd))
        sd = sv.sd_long
        mu = sv.solve_transposed(jz[sd][:, sd], g)
        mu_f = sv.expand(mu, fd).reshape(-1)
        r = hg[:, :3, :3]
        mu_n = mu_f.view(-1, 6)
        gf = -torch.cat(((r @ mu_n[:, :3, None])[..., 0], (r @ mu_n[:, 3:, None])[..., 0]), -1)
        return gf, sv.k_gradient(mu_f, p, eps)

    @staticmethod
    def pullback(sv, hg, fd, ks, g_hg):
        """(dL/df_dead, dL/dk_scale) from dL/dhg at the solution hg."""
        tb = sv.tables
        g = sv.reduce(lie_gradient(hg, g_hg))

        def res(xs):
            xi = sv.expand(xs, fd)
            return sv.reduce(beam_residual(tb, hg @ exp_se3(xi), ks, None, fd))

        mu = sv.solve_transposed(sv.jacobian(res, torch.zeros(sv.n_s, dtype=F64, device=hg.device)), g)
        with torch.enable_grad():
            fd_ = fd.clone().requires_grad_(True)
            ks_ = ks.clone().requires_grad_(True)
            r = sv.reduce(beam_residual(tb, hg, ks_, None, fd_))
            gf, gk = torch.autograd.grad(r, (fd_, ks_), -mu)
        return gf, gk


class DynamicBeamSolve(torch.autograd.Function):
    """The solve-DOF components of phi_a, the generalised-alpha Newton
    solution of one load step from the start values (phi0, v0, v_dot0) at
    the alpha point (structure/beam.py:2170-2316): R = 0 with
    phi_a = phi0 except at the solve DOFs, hg_u = hg_prev exp(phi_a),
    v_a = v0 + gamma' (phi_a - phi0), v_dot_a = v_dot0 + beta' (phi_a - phi0)
    and global dead loads f_dead.  Backward as StaticBeamSolve, with the
    Jacobian over the solve DOFs including the velocity and acceleration
    terms (coupled/gradients/coupled.py:540-629).  ``full_state`` composes
    phi_a, v_a and v_dot_a from the result differentiably."""

    @staticmethod
    def forward(ctx, hg_prev, phi0, v0, vd0, f_dead, k_scale, solver):
        phi = solver.dynamic(hg_prev.detach(), phi0.detach(), v0.detach(), vd0.detach(), f_dead.detach(),
                             k_scale.detach())
        phi_s = solver.reduce(phi)
        ctx.solver = solver
        ctx.save_for_backward(phi_s, hg_prev, phi0, v0, vd0, f_dead, k_scale)
        return phi_s

    @staticmethod
    def full_state(sv, phi_s, phi0, v0, vd0):
        phi = phi0 + sv.expand(phi_s - sv.reduce(phi0), phi0)
        dphi = phi - phi0
        return phi, v0 + sv.ti.gamma_prime * dphi, vd0 + sv.ti.beta_prime * dphi

    @staticmethod
    def residual(sv, phi_s, hg_prev, phi0, v0, vd0, f_dead, k_scale):
        phi, v, vd = DynamicBeamSolve.full_state(sv, phi_s, phi0, v0, vd0)
        return sv.reduce(beam_residual(sv.tables, hg_prev @ exp_se3(phi), k_scale, None, f_dead, v, vd))

    @staticmethod
    def backward(ctx, g_phi_s):
        phi_s, *inputs = ctx.saved_tensors
        sv = ctx.solver
        args = (phi_s.detach(), *[x.detach() for x in inputs], g_phi_s.contiguous())
        fn = DynamicBeamSolve.pullback_device if (sv.device_pullback and phi_s.is_cuda) else DynamicBeamSolve.pullback
        grads = sv.run_backward("dynamic", lambda *x: fn(sv, *x), args)
        return (*grads, None)

    @staticmethod
    def pullback_device(sv, phi_s, hg_prev, phi0, v0, vd0, f_dead, k_scale, g_phi_s):
        """pullback from the exact element derivative blocks of the device
        kernel.  With hg = hg_prev exp(phi) and dR/dzeta the derivative in
        right increments of hg (assembled dfe/dd P plus dead-load blocks):
        J = dR/dzeta T(phi) + gamma' dR/dv - beta' M_t over the solve DOFs;
        mu = J^-T g; then -mu^T dR/d(inputs) through the linear maps of
        full_state, Ad_{exp(-phi)} for hg_prev (returned as a matrix gradient
        exact on every tangent direction hg_prev X, X in se(3), the only
        directions hg_prev moves in), R mu for the dead loads and P^T K eps
        for the stiffness scalings."""
        tb, ti = sv.tables, sv.ti
        n = tb.n_nodes
        phi, v, vd = DynamicBeamSolve.full_state(sv, phi_s, phi0, v0, vd0)
        e_phi = exp_se3(phi)
        hg = hg_prev @ e_phi
        k_elem = (tb.k_cs * k_scale[:, None, None])[tb.k_index]
        _, jd, jv, mt, p, eps = element_jacobians(sv.beam, hg, k_elem, v, vd)
        jz = _node_blocks_add(sv.assemble_blocks(jd @ p), _dead_stiffness(hg, f_dead))
        jv_f = sv.assemble_blocks(jv)
        mt_f = sv.assemble_blocks(mt)
        t = t_se3(phi)
        n6 = 6 * n
        jzt = (jz.view(n6, n, 6).transpose(0, 1).contiguous() @ t).transpose(0, 1).reshape(n6, n6)
        jac = jzt + ti.gamma_prime * jv_f - ti.beta_prime * mt_f
        sd = sv.sd_long
        mu = sv.solve_transposed(jac[sd][:, sd], g_phi_s)
        mu_f = sv.expand(mu, phi0).reshape(-1)
        gz = -(jz.transpose(0, 1) @ mu_f)
        gv = -(jv_f.transpose(0, 1) @ mu_f)
        gvd = mt_f.transpose(0, 1) @ mu_f
        gphi = (t.transpose(-1, -2) @ gz.view(n, 6, 1))[..., 0].reshape(-1)
        sel = sv.solve_selector(phi_s.device)
        g_phi0 = (1.0 - sel) * gphi - sel * (ti.gamma_prime * gv + ti.beta_prime * gvd)
        # hg_prev exp(eta) exp(phi) = hg exp(Ad_{exp(-phi)} eta): g_eta = Ad^T g_zeta
        rp, pp = e_phi[:, :3, :3], e_phi[:, :3, 3]
        gzn = gz.view(n, 6)
        eta_u = (rp @ gzn[:, :3, None])[..., 0]
        pl = (rp.transpose(-1, -2) @ pp[..., None])[..., 0]
        eta_w = (rp @ (gzn[:, 3:] + torch.linalg.cross(pl, gzn[:, :3], dim=-1))[..., None])[..., 0]
        h = torch.nn.functional.pad(torch.cat((0.5 * skew(eta_w), eta_u[..., None]), -1), (0, 0, 0, 1))
        g_hg_prev = hg_inv(hg_prev).transpose(-1, -2) @ h
        r = hg[:, :3, :3]
        mu_n = mu_f.view(n, 6)
        gf = -torch.cat(((r @ mu_n[:, :3, None])[..., 0], (r @ mu_n[:, 3:, None])[..., 0]), -1)
        return (g_hg_prev, g_phi0.view(n, 6), gv.view(n, 6), gvd.view(n, 6), gf, sv.k_gradient(mu_f, p, eps))

    @staticmethod
    def pullback(sv, phi_s, hg_prev, phi0, v0, vd0, f_dead, k_scale, g_phi_s):
        """dL/d(hg_prev, phi0, v0, vd0, f_dead, k_scale) from dL/dphi_s."""
        inputs = (hg_prev, phi0, v0, vd0, f_dead, k_scale)

        def res(xs):
            return DynamicBeamSolve.residual(sv, phi_s + xs, *inputs)

        jac = sv.jacobian(res, torch.zeros(sv.n_s, dtype=F64, device=phi_s.device))
        mu = sv.solve_transposed(jac, g_phi_s)
        with torch.enable_grad():
            leaves = [x.clone().requires_grad_(True) for x in inputs]
            r = DynamicBeamSolve.residual(sv, phi_s, *leaves)
            grads = torch.autograd.grad(r, leaves, -mu, allow_unused=True)
        return grads
