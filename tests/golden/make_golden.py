"""Writes the golden fixtures of the small-uniform workload (n=2,000 train /
500 test, d=3, m=16) from the FP64 oracle.

The oracle itself is pinned by tests/test_oracle.py (scikit-learn exact GP at
m >= n-1, dense conditionals, finite differences, lexsort kNN); these fixtures
freeze its outputs so the GPU tests can compare against stored vectors as well.
Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from nngpr_amd import datasets  # noqa: E402
from oracle import vecchia_oracle as vo  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "small_uniform.npz")

# Free-theta variant of the small-uniform kernel for the gradient fixture:
# theta = log([noise, constant, length_scale]) of White + C * RBF.
THETA_MAP = {"noise_level": 0, "rbf_level": 1, "length_scale": 2}


def main():
    ds = datasets.small_uniform()
    x, y, xq, m = ds["x_train"], ds["y_train"], ds["x_test"], ds["num_nn"]
    w, s, ls, alpha = 0.01, 1.0, 0.3, ds["alpha"]
    nn = vo.vecchia_knn(x, ls, m)
    ll, _, llp = vo.vecchia_loglik(x, y, nn, w, 0.0, s, ls, alpha)
    _, grad, _ = vo.vecchia_loglik(x, y, nn, w, 0.0, s, ls, alpha, THETA_MAP, 3, eval_gradient=True)
    nnq = vo.knn(x, xq, ls, m)
    mean, std = vo.predict(x, y, xq, nnq, w, 0.0, s, ls, alpha)
    np.savez_compressed(OUT, nn_train=nn, nn_test=nnq, loglik=np.array(ll), loglik_points=llp,
                        grad=grad, mean=mean[:, 0], std=std)
    print("wrote", OUT, "loglik", ll)


if __name__ == "__main__":
    main()
