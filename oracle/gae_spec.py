"""ORACLE / TEST INFRASTRUCTURE — CPU restatement (numpy float32, scalar
steps in the reference's order) of the trainer's generalized advantage
estimation, GGL::GAE::Compute (GigaLearnCPP/src/private/GigaLearnCPP/PPO/
GAE.cpp:128-195, the general backward loop; its all-truncated fast path,
:86-126, is the same arithmetic).  Only tests use it; the engine's device GAE
(rsx_gae) never does.

Inputs: rewards (T), terminals (T, int8: 0 none, 1 normal, 2 truncated),
value predictions (T), bootstrap values of the truncated steps in buffer order
(walking backwards, the k-th truncation from the end takes
trunc_val_preds[num_truncs - 1 - k], GAE.cpp:152-162), gamma, lambda,
return_std, clip_range.  A last step that is not a terminal bootstraps from 0
and ends the return chain (GAE.cpp:163-168).

Returns (advantages, target values, returns, reward clip portion)."""
import numpy as np

F = np.float32


def gae(rews, terms, vals, trunc_vals, gamma, lam, return_std=0.0, clip_range=0.0):
    rews = np.asarray(rews, F)
    terms = np.asarray(terms, np.int8)
    vals = np.asarray(vals, F)
    trunc_vals = np.asarray(trunc_vals, F)
    T = len(rews)
    g, l, rs, cr = F(gamma), F(lam), F(return_std), F(clip_range)
    adv = np.zeros(T, F)
    ret = np.zeros(T, F)
    prev_lambda, prev_ret = F(0), F(0)
    total_rew, total_clip = F(0), F(0)
    trunc_count = 0
    n_trunc = len(trunc_vals)
    for step in range(T - 1, -1, -1):
        term = int(terms[step])
        done = F(1) if term == 1 else F(0)
        trunc = F(1) if term == 2 else F(0)
        if rs != 0:
            cur = F(rews[step] / rs)
            total_rew = F(total_rew + abs(cur))
            if cr > 0:
                cur = F(min(max(cur, -cr), cr))
            total_clip = F(total_clip + abs(cur))
        else:
            cur = rews[step]
            total_rew = F(total_rew + abs(cur))
        if term == 2:
            if trunc_count >= n_trunc:
                raise ValueError("more truncated terminals than bootstrap values")
            nxt = trunc_vals[n_trunc - 1 - trunc_count]
            trunc_count += 1
        elif step >= T - 1:
            nxt = F(0)
            trunc = F(1)
        else:
            nxt = vals[step + 1]
        pred_return = F(cur + F(F(g * nxt) * F(F(1) - done)))
        delta = F(pred_return - vals[step])
        cur_ret = F(rews[step] + F(F(F(prev_ret * g) * F(F(1) - done)) * F(F(1) - trunc)))
        ret[step] = cur_ret
        prev_lambda = F(delta + F(F(F(F(g * l) * F(F(1) - done)) * F(F(1) - trunc)) * prev_lambda))
        adv[step] = prev_lambda
        prev_ret = cur_ret
    if trunc_count != n_trunc:
        raise ValueError("truncated terminal count does not match the bootstrap values")
    target = (vals + adv).astype(F)
    clip_portion = F(F(total_rew - total_clip) / max(total_rew, F(1e-7)))
    return adv, target, ret, clip_portion
