"""NUTS with the RMHMC integrator for a general dense user metric
(geomjax/rmhmc/nuts.py with any ``metric_fn``).

The iterative doubling of geomjax/mcmc/trajectory.py:149-317,517-668 with the
numpyro U-turn checks of geomjax/mcmc/termination.py:30-119 and the progressive
sampling of geomjax/mcmc/proposal.py:198-254, batched over chains in lock-step:
at expansion e every chain still expanding builds a subtree of 2^e steps in its
own direction, a chain whose subtree turned or diverged is masked for the rest
of it, and a chain whose trajectory stopped is masked for the rest of the
transition (the per-chain index of its progressive-sampling uniforms advances
only with its own steps).  The steps are the dense implicit midpoint of
``rmhmc_dense`` (torch.func derivatives of the user's functions,
``gjx_dense_rmhmc_field`` for the metric algebra).  float64, D <= 64.
"""
import math

import torch

from .lmcmc_dense import _check, _metric, _value_and_grad
from .rmhmc_dense import _field, _implicit_midpoint_step, _kinetic

__all__ = ["transition", "expand", "lmc_transition"]

_CRITERIA = ("euc", "bet", "riem")


def _sel(mask, a, b):
    """Per-chain select over tuples of [C, ...] tensors."""
    if isinstance(a, tuple):
        return tuple(_sel(mask, x, y) for x, y in zip(a, b))
    m = mask.reshape(mask.shape + (1,) * (a.dim() - 1))
    return torch.where(m, a, b)


def _turning(criterion, m_l, m_r, m_sum, v_l, v_r, v_sum):
    """geomjax/rmhmc/metrics.py:100-117 per chain."""
    if criterion == "euc":
        rho, left, right = v_sum, v_l, v_r
    elif criterion == "bet":
        rho, left, right = m_sum, v_l, v_r
    else:
        rho, left, right = v_sum, m_l, m_r
    return ((left * rho).sum(-1) <= 0) | ((right * rho).sum(-1) <= 0)


def _ckpt_idxs(n):
    """geomjax/mcmc/termination.py:71-87."""
    idx_max = bin(n >> 1).count("1")
    ones, m = 0, n
    while m & 1:
        ones += 1
        m >>= 1
    return idx_max - ones + 1, idx_max


def transition(noise, state, logdensity_fn, step_size, metric_fn, solver, max_num_doublings=10,
               stopping_criterion="euc", divergence_threshold=1000.0):
    """One NUTS transition of every chain with the RMHMC implicit midpoint.
    ``noise`` = (z [C, D], u_direction [C, M], u_merge [C, M], u_step [C, 2^M - 1]).
    Returns (position, logdensity, grad, info dict)."""
    from .lmcmc_dense import resolve_metric_fn

    metric_fn = resolve_metric_fn(logdensity_fn, metric_fn)
    q0 = state.position.contiguous()
    _check(q0)
    C, D = q0.shape
    z = noise[0]
    l0 = state.logdensity.reshape(C).to(torch.float64)
    g0 = state.logdensity_grad.reshape(C, D).to(torch.float64).contiguous()
    G, _ = _metric(metric_fn, q0)
    p0, v0, _, logdet0, pu0 = _field(G, None, z=z)
    H0 = -l0 + _kinetic(logdet0, pu0, D)

    def step(cur, eps_vec):
        qn, pn, _ = _implicit_midpoint_step(cur[0], cur[1], eps_vec, logdensity_fn, metric_fn, solver)
        ln, gn = _value_and_grad(logdensity_fn, qn)
        Gq, _ = _metric(metric_fn, qn)
        _, vn, _, logdet, pu = _field(Gq, None, p=pn)
        return (qn, pn, vn, ln, gn), -ln + _kinetic(logdet, pu, D)

    P_s, info = expand((q0, p0, v0, l0, g0), H0, step, float(step_size), noise, max_num_doublings,
                       stopping_criterion, divergence_threshold)
    info["momentum"] = p0
    return P_s[0], P_s[3], P_s[4], info


def expand(S0, H0, step, eps, noise, max_num_doublings, stopping_criterion, divergence_threshold):
    """The batched iterative expansion from the state tuple S0 = (position,
    momentum, velocity, ...) with energies H0 [C]; ``step(cur, signed_eps [C])``
    returns (new tuple, its energy).  Returns (proposal tuple, info dict)."""
    if stopping_criterion not in _CRITERIA:
        raise ValueError(f"stopping_criterion must be one of {list(_CRITERIA)} (got {stopping_criterion!r})")
    q0, p0, v0 = S0[0], S0[1], S0[2]
    C, D = q0.shape
    dev = q0.device
    M = int(max_num_doublings)
    _, u_dir, u_merge, u_step = noise

    def weight(E):
        w = H0 - E
        return torch.where(torch.isnan(w), torch.full_like(w, -math.inf), w)

    ninf = torch.full((C,), -math.inf, dtype=torch.float64, device=dev)
    zeros = torch.zeros(C, dtype=torch.float64, device=dev)
    # proposal: (state, energy, weight, sum_log_p_accept)
    P_s, P_e, P_w, P_a = S0, H0.clone(), zeros.clone(), ninf.clone()
    left, right = S0, S0
    msum, vsum = p0.clone(), v0.clone()
    nstates = torch.zeros(C, dtype=torch.int64, device=dev)
    r_ck = torch.zeros(C, M, D, dtype=torch.float64, device=dev)
    rs_ck, v_ck, vs_ck = torch.zeros_like(r_ck), torch.zeros_like(r_ck), torch.zeros_like(r_ck)
    j = torch.zeros(C, dtype=torch.int64, device=dev)
    done = torch.zeros(C, dtype=torch.bool, device=dev)
    diverging = torch.zeros_like(done)
    turning = torch.zeros_like(done)
    expansions = torch.zeros(C, dtype=torch.int64, device=dev)
    capturing = torch.cuda.is_current_stream_capturing()
    for e in range(M):
        active = ~done
        if not capturing and not bool(active.any()):
            break
        fwd = u_dir[:, e] < 0.5
        sgn = torch.where(fwd, 1.0, -1.0).to(torch.float64)
        start = _sel(fwd, right, left)
        # the subtree (dynamic_progressive_integration, trajectory.py:149-317)
        sub_l = sub_r = start
        s_msum, s_vsum = start[1].clone(), start[2].clone()
        s_n = torch.zeros_like(nstates)
        sP_s, sP_e, sP_w, sP_a = start, zeros.clone(), ninf.clone(), ninf.clone()
        cur = start
        sub_active = active.clone()
        div_s = torch.zeros_like(done)
        term_s = torch.zeros_like(done)
        for k in range(2 ** e):
            if not capturing and not bool(sub_active.any()):
                break
            new, En = step(cur, sgn * eps)
            pn, vn = new[1], new[2]
            wn = weight(En)
            an = torch.minimum(wn, zeros)
            div_now = -wn > divergence_threshold
            if k == 0:
                n_l, n_r, n_m, n_v, n_n = new, new, pn.clone(), vn.clone(), torch.ones_like(s_n)
                nP = (new, En, wn, an)
            else:
                n_l, n_r, n_m, n_v, n_n = sub_l, new, s_msum + pn, s_vsum + vn, s_n + 1
                # progressive_uniform_sampling (proposal.py:198-222)
                dw = wn - sP_w
                pr = torch.where(dw > -700, torch.sigmoid(dw), zeros)
                u = u_step.gather(1, j.clamp(max=u_step.shape[1] - 1)[:, None])[:, 0]
                take = u < pr
                nP = (_sel(take, new, sP_s), torch.where(take, En, sP_e), torch.logaddexp(sP_w, wn),
                      torch.logaddexp(sP_a, an))
            sa = sub_active
            sub_l, sub_r = _sel(sa, n_l, sub_l), _sel(sa, n_r, sub_r)
            s_msum, s_vsum, s_n = _sel(sa, n_m, s_msum), _sel(sa, n_v, s_vsum), torch.where(sa, n_n, s_n)
            sP_s = _sel(sa, nP[0], sP_s)
            sP_e, sP_w, sP_a = (torch.where(sa, x, y) for x, y in ((nP[1], sP_e), (nP[2], sP_w), (nP[3], sP_a)))
            j = j + sa.to(torch.int64)
            # iterative U-turn checks (termination.py:30-69, 89-117)
            i_min, i_max = _ckpt_idxs(k)
            if k % 2 == 0:
                m3 = sa[:, None]
                r_ck[:, i_max] = torch.where(m3, pn, r_ck[:, i_max])
                rs_ck[:, i_max] = torch.where(m3, s_msum, rs_ck[:, i_max])
                v_ck[:, i_max] = torch.where(m3, vn, v_ck[:, i_max])
                vs_ck[:, i_max] = torch.where(m3, s_vsum, vs_ck[:, i_max])
            term_now = torch.zeros_like(done)
            for i in range(i_max, i_min - 1, -1):
                term_now = term_now | _turning(stopping_criterion, r_ck[:, i], pn, s_msum - rs_ck[:, i] + r_ck[:, i],
                                               v_ck[:, i], vn, s_vsum - vs_ck[:, i] + v_ck[:, i])
            term_s = term_s | (sa & term_now)
            div_s = div_s | (sa & div_now)
            cur = _sel(sa, new, cur)
            sub_active = sa & ~(term_now | div_now)
        # the subtree's ends in trajectory order
        s_left, s_right = _sel(fwd, sub_l, sub_r), _sel(fwd, sub_r, sub_l)
        # merge (trajectory.py:580-660): biased progressive sampling unless the subtree turned or diverged
        stop = div_s | term_s
        dw = sP_w - P_w
        pb = torch.where(dw < 700, torch.clamp(torch.exp(dw), max=1.0), torch.ones_like(dw))
        take = (~stop) & (u_merge[:, e] < pb)
        mP_s = _sel(take, sP_s, P_s)
        mP_e = torch.where(take, sP_e, P_e)
        mP_w = torch.where(stop, P_w, torch.logaddexp(P_w, sP_w))
        mP_a = torch.logaddexp(P_a, sP_a)
        t_left, t_right = _sel(fwd, left, s_left), _sel(fwd, s_right, right)
        t_msum, t_vsum, t_n = msum + s_msum, vsum + s_vsum, nstates + s_n
        turn_full = _turning(stopping_criterion, t_left[1], t_right[1], t_msum, t_left[2], t_right[2], t_vsum)
        a = active
        P_s = _sel(a, mP_s, P_s)
        P_e, P_w, P_a = (torch.where(a, x, y) for x, y in ((mP_e, P_e), (mP_w, P_w), (mP_a, P_a)))
        left, right = _sel(a, t_left, left), _sel(a, t_right, right)
        msum, vsum, nstates = _sel(a, t_msum, msum), _sel(a, t_vsum, vsum), torch.where(a, t_n, nstates)
        expansions = expansions + a.to(torch.int64)
        diverging = torch.where(a, div_s, diverging)
        turning = torch.where(a, term_s | turn_full, turning)
        done = done | (a & (div_s | term_s | turn_full))
    acceptance = torch.exp(P_a) / nstates.to(torch.float64)
    info = dict(is_divergent=diverging, is_turning=turning, energy=P_e, left=left[0], right=right[0],
                num_trajectory_expansions=expansions.to(torch.int32), num_integration_steps=nstates.to(torch.int32),
                acceptance_rate=acceptance)
    return P_s, info


def lmc_transition(noise, state, logdensity_fn, step_size, metric_fn, max_num_doublings=10, stopping_criterion="euc",
                   divergence_threshold=1000.0):
    """One NUTS transition with the general-metric LMC integrator
    (geomjax/lmcmc/nuts.py:78-161): velocity v0 = L^-T z, momentum G v0, the
    energy -l - log det G / 2 + v^T G v / 2 - J carried along the trajectory.
    Returns (position, logdensity, grad, volume_adjustment, info dict)."""
    from .lmcmc_dense import _position, lmc_step, resolve_metric_fn

    metric_fn = resolve_metric_fn(logdensity_fn, metric_fn)

    q0 = state.position.contiguous()
    _check(q0)
    C, D = q0.shape
    dev = q0.device
    z = noise[0]
    l0 = state.logdensity.reshape(C).to(torch.float64)
    g0 = state.logdensity_grad.reshape(C, D).to(torch.float64).contiguous()
    J0 = torch.as_tensor(state.volume_adjustment, dtype=torch.float64, device=dev).expand(C).clone()
    G, dG = _metric(metric_fn, q0)
    logdet0, _, v0, _ = _position(G, dG, z=z)
    _, _, _, vgv0 = _position(G, dG, velocity=v0)
    m0 = torch.bmm(G, v0[:, :, None])[:, :, 0]
    H0 = -l0 + (-0.5 * logdet0 + 0.5 * vgv0) - J0

    def step(cur, eps_vec):
        return lmc_step(cur, eps_vec, logdensity_fn, metric_fn)

    P_s, info = expand((q0, m0, v0, l0, g0, J0), H0, step, float(step_size), noise, max_num_doublings,
                       stopping_criterion, divergence_threshold)
    info["momentum"] = v0  # NUTSInfo's first field holds the initial velocity (lmcmc/nuts.py:36-77)
    return P_s[0], P_s[3], P_s[4], P_s[5], info
