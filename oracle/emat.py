"""Face-coupling matrices, restated branch by branch.

``E_space`` follows ``src/rt/get_Ex.py:10-127`` (and the identical
``get_Ey.py:9-125``): E in R^{n1 x n0} couples the trace basis of the
neighbour column (index i, order n1) to the own column (index p, order n0)
along the shared edge.  ``E_theta`` follows ``src/rt/get_Eth.py:9-150`` with
the outward-direction factor Theta_F(th) = cos(th - F pi / 2) folded into the
quadrature weights and entries below 1e-15 zeroed (``get_Eth.py:149``).

Branches (own = 0, neighbour = 1, dlv = lv0 - lv1):
  dlv in {0, +1}, n0 >= n1 : E[i,p] = w0_p l1_i(T01 xi0_p)
  dlv in {0, +1}, n0 <  n1 : E[i,p] = sum_i' w1_i' l1_i(T01 xi1_i') l0_p(xi1_i')
  dlv = -1,       n0 >  n1 : E[i,p] = 1/2 sum_p' w0_p' l1_i(xi0_p') l0_p(T10 xi0_p')
  dlv = -1,       n0 <= n1 : E[i,p] = 1/2 w1_i l0_p(T10 xi1_i)
with T01 = pull_back_1 o push_forward_0 and T10 = pull_back_0 o push_forward_1.

Pins: closed forms with explicit affine maps (xi +- 1)/2 in
tests/test_oracle_pins.py, the identity-reduction of the second branch,
partition of unity (sum_i l1_i = 1) giving exact face integrals, and the
same-level equal-order case E = diag(w).
"""
import numpy as np

from .quad import lgl, lg, lag, push, pull


def _rel(lv0, mid0, lv1, mid1):
    return lv0 - lv1


def E_space(n0, a0, b0, lv0, n1, a1, b1, lv1, rule=lgl):
    x0b, w0 = rule(n0)
    x1b, w1 = rule(n1)
    E = np.zeros((n1, n0))
    dlv = lv0 - lv1
    if dlv not in (-1, 0, 1):
        raise ValueError("level difference outside {-1, 0, 1}")
    if dlv == -1:
        c = 0.5
        if n0 > n1:
            t = pull(a0, b0, push(a1, b1, x0b))
            for i in range(n1):
                for p in range(n0):
                    s = 0.0
                    for pp in range(n0):
                        s += c * w0[pp] * lag(x1b, i, x0b[pp]) * lag(x0b, p, t[pp])
                    E[i, p] = s
        else:
            t = pull(a0, b0, push(a1, b1, x1b))
            for i in range(n1):
                for p in range(n0):
                    E[i, p] = c * w1[i] * lag(x0b, p, t[i])
    else:
        if n0 >= n1:
            t = pull(a1, b1, push(a0, b0, x0b))
            for i in range(n1):
                for p in range(n0):
                    E[i, p] = w0[p] * lag(x1b, i, t[p])
        else:
            t = pull(a1, b1, push(a0, b0, x1b))
            for i in range(n1):
                for p in range(n0):
                    s = 0.0
                    for ii in range(n1):
                        s += w1[ii] * lag(x1b, i, t[ii]) * lag(x0b, p, x1b[ii])
                    E[i, p] = s
    return E


def theta_F(th, F):
    return np.cos(np.asarray(th) - F * np.pi / 2.0)


def E_theta(n0, a0, b0, lv0, n1, a1, b1, lv1, F):
    t0b, w0 = lg(n0)
    t1b, w1 = lg(n1)
    E = np.zeros((n1, n0))
    dlv = lv0 - lv1
    if dlv not in (-1, 0, 1):
        raise ValueError("level difference outside {-1, 0, 1}")
    if dlv == -1:
        if n0 > n1:
            thf = push(a1, b1, t0b)           # own nodes placed in the child
            TF = theta_F(thf, F)
            t = pull(a0, b0, thf)
            for a in range(n1):
                for r in range(n0):
                    s = 0.0
                    for rr in range(n0):
                        s += w0[rr] * TF[rr] * lag(t1b, a, t0b[rr]) * lag(t0b, r, t[rr])
                    E[a, r] = s
        else:
            thf = push(a1, b1, t1b)
            TF = theta_F(thf, F)
            t = pull(a0, b0, thf)
            for a in range(n1):
                for r in range(n0):
                    E[a, r] = w1[a] * TF[a] * lag(t0b, r, t[a])
        E *= 0.5
    else:
        if n0 >= n1:
            thf = push(a0, b0, t0b)
            TF = theta_F(thf, F)
            t = pull(a1, b1, thf)
            for a in range(n1):
                for r in range(n0):
                    E[a, r] = w0[r] * TF[r] * lag(t1b, a, t[r])
        else:
            thf = push(a0, b0, t1b)
            TF = theta_F(thf, F)
            t = pull(a1, b1, thf)
            for a in range(n1):
                for r in range(n0):
                    s = 0.0
                    for aa in range(n1):
                        s += w1[aa] * TF[aa] * lag(t1b, a, t[aa]) * lag(t0b, r, t1b[aa])
                    E[a, r] = s
    E[np.abs(E) < 1.0e-15] = 0.0
    return E
