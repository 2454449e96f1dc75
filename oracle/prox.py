"""Oracle: proximal steps of the PDFB iteration, float64.

Test infrastructure only (see ``oracle/__init__.py``).  Each function cites
the reference file:line whose semantics it restates.
"""
import numpy as np

MATLAB_EPS = np.finfo(np.float64).eps


def proj_l2ball(x, eps, y):
    """``src/proj_l2ball.m:20-22``: p = y + (x - y) * min(eps/||x - y||, 1)."""
    p = x - y
    nrm = np.linalg.norm(p.ravel())
    scale = min(eps / nrm, 1.0) if nrm > 0 else 1.0
    return p * scale + y


def sc(z, radius):
    """Moreau form used live: ``src/update_dual_data_fidelity.m:83``.

    ``z * (1 - min(radius/||z||, 1))``.
    """
    nrm = np.linalg.norm(z.ravel())
    if nrm == 0:
        return z * 0.0
    return z * (1.0 - min(radius / nrm, 1.0))


def solver_proj_elipse_fb(u_scaled, r, y, pU, epsilon, z0, max_itr, min_itr, tol):
    """Projection onto B(y, epsilon) in the metric diag(pU), forward-backward.

    Called at ``src/update_dual_data_fidelity.m:133-135`` as
    ``solver_proj_elipse_fb(v2./pU, r2, y, pU, epsilon, proj, max, min, eps)``
    (submodule RI-measurement-operator, absent).  Restated: with
    ``w = u_scaled + r`` the step solves
    ``argmin_{||z - y|| <= eps} (z - w)^H diag(pU) (z - w)`` by projected
    gradient with step ``1/max(pU)``, warm-started at ``z0``; it stops after
    ``min_itr`` iterations once ``||z_k - z_{k-1}|| <= tol ||z_k||`` or at
    ``max_itr``.  Parity unpinned: the submodule's exact iterates cannot be
    checked; pinned by the pU = const closed form (equals ``proj_l2ball``),
    feasibility and the fixed-point residual.
    """
    w = u_scaled + r
    mu = 1.0 / np.max(pU)
    z = z0.copy()
    it = 0
    for it in range(1, max_itr + 1):
        z_old = z
        g = z - mu * pU * (z - w)
        z = proj_l2ball(g, epsilon, y)
        nz = np.linalg.norm(z)
        crit = np.linalg.norm(z - z_old) / nz if nz > 0 else 0.0
        if it >= min_itr and crit <= tol:
            break
    return z, it


def soft_l21(w, thresh):
    """Weighted l21 soft threshold, Moreau form of
    ``src/fhs/fhs_update_dual_sparsity.m:84-87`` /
    ``src/hs/hs_update_dual_sparsity_distributed.m:10-12``.

    ``w``: (s, L); ``thresh``: (s,) = beta1 * weights1.
    Returns ``w - max(l2 - thresh, 0)/(l2 + eps) * w`` with l2 the row norms.
    """
    l2 = np.sqrt(np.sum(np.abs(w) ** 2, axis=1))
    shrink = np.maximum(l2 - thresh, 0.0) / (l2 + MATLAB_EPS)
    return w - shrink[:, None] * w


def soft_l1_dual(r, thresh):
    """SARA per-basis dual update ``src/sara/sara.m:925``:
    ``r - sign(r) .* max(|r| - thresh, 0)``."""
    return r - np.sign(r) * np.maximum(np.abs(r) - thresh, 0.0)


def nuclear_prox_dual(Z, thresh):
    """``src/fhs/fhs_update_dual_lowrankness.m:41-44`` (and
    ``src/hs/hs_update_dual_lowrankness_serial.m:7-8``):
    ``Z - U diag(max(S - thresh, 0)) V^T`` with the economy SVD of Z (P x L);
    ``thresh`` pairs with the singular values in descending order."""
    U, S, Vt = np.linalg.svd(Z, full_matrices=False)
    return Z - (U * np.maximum(S - thresh, 0.0)) @ Vt


def nuclear_norm(x):
    """``src/nuclear_norm.m:25-26``: sum of singular values of reshape(x, N, L)."""
    L = x.shape[-1]
    X = np.reshape(x, (-1, L), order='F')
    return float(np.sum(np.abs(np.linalg.svd(X, compute_uv=False))))


def update_primal(xsol, g):
    """``src/fhs/fhs_update_primal.m:33-38``."""
    prev = xsol
    x = np.maximum(np.real(xsol - g), 0.0)
    xhat = 2 * x - prev
    rel_x = float(np.sum((prev - x) ** 2))
    norm_x = float(np.sum(x ** 2))
    return x, xhat, rel_x, norm_x


def update_epsilon(epsilon, t, t_block, norm_res, tol_in, tol_out, steps, change):
    """``src/update_epsilon.m:52-71`` (cells as nested lists; returns copies)."""
    epsilon = [list(e) for e in epsilon]
    t_block = [list(tb) for tb in t_block]
    for i in range(len(epsilon)):
        for j in range(len(epsilon[i])):
            if t > t_block[i][j] + steps:
                if norm_res[i][j] < tol_in * epsilon[i][j]:
                    epsilon[i][j] = change * norm_res[i][j] + (1 - change) * epsilon[i][j]
                    t_block[i][j] = t
                if norm_res[i][j] > tol_out * epsilon[i][j]:
                    epsilon[i][j] = change * norm_res[i][j] + (1 - change) * epsilon[i][j]
                    t_block[i][j] = t
    return epsilon, t_block
