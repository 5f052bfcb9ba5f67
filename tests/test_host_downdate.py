"""The plain path's QR downdate (householder_downdate in csrc/plain.hip, the
qr!(R[:, 2:m]) of aa_qr_update!, src/utils/anderson.jl:665-703) compiled for the
host with g++ and checked against what the factorisation must satisfy: R[:, 1:m] =
Qx R', Qx with orthonormal columns, R' upper triangular, the last column of R
cleared; and, for the same input, the same bits as a plain-loop form of the same
Householder steps.  No GPU."""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
PLAIN = os.path.join(HERE, "..", "cdft.jl_amd", "csrc", "plain.hip")

# The same Householder steps with loops over the live ranges only (LAPACK geqrf
# conventions, beta = -sign(alpha) hypot(alpha, ||x||)); the unrolled device form must
# give its bits exactly.
LOOPS = r"""
static void hh_loops(double *R, int m, double *Qx)
{
    const int MX = CDFT_MAXM, nr = m, ncol = m - 1;
    double A[CDFT_MAXM * CDFT_MAXM], tau[CDFT_MAXM], v[CDFT_MAXM];
    for (int j = 0; j < ncol; j++)
        for (int i = 0; i < nr; i++) A[j * MX + i] = R[(j + 1) * m + i];
    for (int j = 0; j < ncol; j++) {
        double alpha = A[j * MX + j], xn = 0.0;
        for (int i = j + 1; i < nr; i++) xn += A[j * MX + i] * A[j * MX + i];
        xn = sqrt(xn);
        tau[j] = 0.0;
        if (xn == 0.0) continue;
        double beta = -copysign(hypot(alpha, xn), alpha);
        tau[j] = (beta - alpha) / beta;
        double sc = 1.0 / (alpha - beta);
        for (int i = j + 1; i < nr; i++) A[j * MX + i] *= sc;
        A[j * MX + j] = beta;
        for (int jj = j + 1; jj < ncol; jj++) {
            double w = A[jj * MX + j];
            for (int i = j + 1; i < nr; i++) w += A[j * MX + i] * A[jj * MX + i];
            w *= tau[j];
            A[jj * MX + j] -= w;
            for (int i = j + 1; i < nr; i++) A[jj * MX + i] -= w * A[j * MX + i];
        }
    }
    for (int j = 0; j < ncol; j++)
        for (int i = 0; i < nr; i++) Qx[j * nr + i] = (i == j) ? 1.0 : 0.0;
    for (int h = ncol - 1; h >= 0; h--) {
        if (tau[h] == 0.0) continue;
        v[h] = 1.0;
        for (int i = h + 1; i < nr; i++) v[i] = A[h * MX + i];
        for (int j = 0; j < ncol; j++) {
            double w = 0.0;
            for (int i = h; i < nr; i++) w += v[i] * Qx[j * nr + i];
            w *= tau[h];
            for (int i = h; i < nr; i++) Qx[j * nr + i] -= w * v[i];
        }
    }
    for (int j = 0; j < ncol; j++)
        for (int i = 0; i < ncol; i++) R[j * m + i] = (i <= j) ? A[j * MX + i] : 0.0;
    for (int i = 0; i < m; i++) R[(m - 1) * m + i] = 0.0;
}
"""

MAIN = r"""
#include <cstdio>
int main()
{
    int m;
    while (scanf("%d", &m) == 1) {
        double R[64], R2[64], Q[64], Q2[64];
        for (int k = 0; k < m * m; k++) { if (scanf("%lf", &R[k]) != 1) return 1; R2[k] = R[k]; }
        for (int k = 0; k < 64; k++) Q[k] = Q2[k] = 0.0;
        householder_downdate(R, m, Q);
        hh_loops(R2, m, Q2);
        for (int k = 0; k < m * m; k++) printf("%a ", R[k]);
        for (int k = 0; k < m * (m - 1); k++) printf("%a ", Q[k]);
        for (int k = 0; k < m * m; k++) printf("%a ", R2[k]);
        for (int k = 0; k < m * (m - 1); k++) printf("%a ", Q2[k]);
        printf("\n");
    }
    return 0;
}
"""


@pytest.fixture(scope="module")
def downdate_bin(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = open(PLAIN).read()
    i0 = src.index("__host__ __device__ inline void householder_downdate(double *R, int m, double *Qx)")
    i1 = src.index("\n}\n", i0) + 3
    fn = src[i0:i1]
    d = tmp_path_factory.mktemp("downdate")
    cpp = d / "dd.cpp"
    cpp.write_text("#include <cmath>\n#define CDFT_MAXM 8\n#define __host__\n#define __device__\n"
                   "using std::sqrt; using std::hypot; using std::copysign;\n" + fn + LOOPS + MAIN)
    exe = d / "dd"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-std=c++17", str(cpp), "-o", str(exe)], check=True)
    return str(exe)


def _cases(seed=3):
    rng = np.random.default_rng(seed)
    out = []
    for m in range(2, 9):
        for t in range(6):
            R = np.triu(rng.normal(size=(m, m)) * 10.0 ** rng.integers(-3, 4))
            if t == 5:
                R[:, 2 % m] = 0.0                        # a zero column
            out.append((m, R))
    return out


def test_downdate_factorisation_and_bits(downdate_bin):
    cases = _cases()
    # column-major storage, as on the device: R[j * m + i] = R[i, j]
    text = "".join(f"{m} " + " ".join(repr(float(x)) for x in R.T.reshape(-1)) + "\n" for m, R in cases)
    res = subprocess.run([downdate_bin], input=text, capture_output=True, text=True, check=True).stdout.splitlines()
    assert len(res) == len(cases)
    for (m, R), line in zip(cases, res):
        v = [float.fromhex(x) for x in line.split()]
        n1, n2 = m * m, m * (m - 1)
        Rd = np.array(v[:n1]).reshape(m, m).T
        Qx = np.array(v[n1:n1 + n2]).reshape(m - 1, m).T          # Qx[j * m + i] = Qx[i, j]
        assert v[n1 + n2:n1 + n2 + n1] == v[:n1] and v[2 * n1 + n2:] == v[n1:n1 + n2], m
        Rp = Rd[:m - 1, :m - 1]
        assert np.all(np.tril(Rp, -1) == 0.0)
        assert np.all(Rd[:, m - 1] == 0.0) and np.all(Rd[m - 1, :] == 0.0)
        A = R[:, 1:]
        scale = max(np.abs(A).max(), 1e-300)
        assert np.allclose(Qx.T @ Qx, np.eye(m - 1), atol=1e-13)
        assert np.allclose(Qx @ Rp, A, atol=1e-12 * scale, rtol=0)
        # LAPACK conventions: R's diagonal is minus the sign of the leading entry of each
        # reflected column (beta = -sign(alpha) hypot(alpha, ||x||)), |R'| as numpy's
        Rn = np.linalg.qr(A, mode="r")
        assert np.allclose(np.abs(np.diag(Rp)), np.abs(np.diag(Rn)), atol=1e-12 * scale, rtol=1e-12)
