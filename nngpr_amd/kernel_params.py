"""Reduction of a kernel tree to the accelerated form  White + C + s * E,  with E
an RBF or a Matern (nu = 1/2, 3/2, 5/2) stationary part.

Restates the reference's RbfAccelerator.get_kernel_params and
get_kernel_theta_mapping (cuda_kernel_accelerators.py:520-649) and the
CudaKernelAccelerator cache (:18-164), with one deliberate difference: Matern is a
subclass of RBF in scikit-learn, and the reference's subclass test lets it through
as if it were an RBF (its accelerated fill then computes RBF entries).  Here a
Matern with nu in {0.5, 1.5, 2.5} keeps its own covariance (the closed forms of
batched_kernels.py:1003-1131, the reference's generic path), nu = inf is the RBF,
and any other nu raises NotImplementedError like every non-accelerated kernel.
"""
import copy
import numbers

import numpy as np
from sklearn.gaussian_process import kernels


def get_kernel_params(kernel):
    """Dict of {noise_level, constant_value, rbf_level, length_scale[, matern_nu]}
    describing the kernel as  noise*[i==j] + constant + rbf_level*E(length_scale),
    E the RBF or, with matern_nu, the Matern of that nu.

    Raises NotImplementedError for kernels outside that family
    (cuda_kernel_accelerators.py:534-590).
    """
    if isinstance(kernel, kernels.Matern):
        nu = float(kernel.nu)
        if not (np.isinf(nu) or nu in (0.5, 1.5, 2.5)):
            raise NotImplementedError("Matern nu=%r is not accelerated (0.5, 1.5, 2.5 or inf)"
                                      % kernel.nu)
        ls = kernel.length_scale
        out = {"length_scale": ls if isinstance(ls, numbers.Number)
               else np.array(ls).reshape((1, -1))}
        if not np.isinf(nu):
            out["matern_nu"] = nu
        return out
    if isinstance(kernel, kernels.WhiteKernel):
        return {"noise_level": kernel.noise_level}
    if isinstance(kernel, kernels.ConstantKernel):
        return {"constant_value": kernel.constant_value}
    if isinstance(kernel, kernels.RBF):
        ls = kernel.length_scale
        return {"length_scale": ls if isinstance(ls, numbers.Number)
                else np.array(ls).reshape((1, -1))}
    if isinstance(kernel, kernels.Sum):
        p1 = get_kernel_params(kernel.k1)
        p2 = get_kernel_params(kernel.k2)
        rbf1 = len(set(p1).difference({"noise_level", "constant_value"})) > 0
        rbf2 = len(set(p2).difference({"noise_level", "constant_value"})) > 0
        if rbf1 and rbf2:
            raise NotImplementedError("Sum of two RBF-type kernels is not supported: %s + %s"
                                      % (p1, p2))
        if rbf1:
            p1, p2 = p2, p1
        if "noise_level" in p1 or "noise_level" in p2:
            p2["noise_level"] = p1.get("noise_level", 0) + p2.get("noise_level", 0)
        if "constant_value" in p1 or "constant_value" in p2:
            p2["constant_value"] = p1.get("constant_value", 0) + p2.get("constant_value", 0)
        return p2
    if isinstance(kernel, kernels.Product):
        p1 = get_kernel_params(kernel.k1)
        p2 = get_kernel_params(kernel.k2)
        c1 = len(set(p1).difference({"constant_value"})) == 0
        c2 = len(set(p2).difference({"constant_value"})) == 0
        if not c1 and not c2:
            raise NotImplementedError("Product is supported only with a ConstantKernel: %s * %s"
                                      % (p1, p2))
        if c2:
            p1, p2 = p2, p1
        if "noise_level" in p2:
            p2["noise_level"] = p1["constant_value"] * p2["noise_level"]
        if "constant_value" in p2:
            p2["constant_value"] = p1["constant_value"] * p2["constant_value"]
        if "rbf_level" in p2 or "length_scale" in p2:
            p2["rbf_level"] = p1["constant_value"] * p2.get("rbf_level", 1)
        return p2
    raise NotImplementedError("Unsupported kernel type: " + str(type(kernel)))


def get_kernel_theta_mapping(kernel):
    """Index in kernel.theta of each parameter of get_kernel_params.

    Clones the kernel with theta = arange(n) and matches every parameter against
    exp(i) within a relative 1e-15 (cuda_kernel_accelerators.py:610-649).  Every
    theta must be matched, ARD indices must be contiguous.
    """
    n_theta = len(kernel.theta)
    epsi = 1e-15
    probe = kernel.clone_with_theta(np.arange(n_theta, dtype=np.float64))
    expected = np.exp(probe.theta)
    params = get_kernel_params(probe)

    def find(name):
        val = params[name]
        if name != "length_scale" or isinstance(val, numbers.Number):
            found = (val < expected * (1 + epsi)) & (val > expected * (1 - epsi))
            nf = int(np.sum(found))
            if nf > 1:
                raise NotImplementedError("Cannot find theta mapping for parameter " + name)
            if nf == 0:
                return None
            return int(np.where(found)[0][0])
        vals = np.asarray(val).reshape(-1)
        found = np.zeros(len(expected), dtype=bool)
        for v in vals:
            hit = (v < expected * (1 + epsi)) & (v > expected * (1 - epsi))
            if np.sum(hit) == 1:
                found[np.where(hit)[0][0]] = True
        nf = int(np.sum(found))
        if nf > vals.size:
            raise NotImplementedError("Cannot find theta mapping for parameter " + name)
        if nf == 0:
            return None
        ind = np.where(found)[0]
        if len(ind) > 1 and set(np.unique(np.diff(ind))) != {1}:
            raise NotImplementedError("Unsupported case of non-adjacent length-scale thetas")
        return ind

    # matern_nu is a fixed property of the kernel, not a hyperparameter
    mapping = {p: find(p) for p in params if p != "matern_nu"}
    mapping = {k: v for k, v in mapping.items() if v is not None}
    n_found = sum(len(v) if isinstance(v, np.ndarray) else 1 for v in mapping.values())
    if n_found != n_theta:
        raise NotImplementedError("Cannot find the mapping for all kernel thetas")
    return mapping


def can_accelerate_kernel(kernel):
    try:
        get_kernel_params(kernel)
        get_kernel_theta_mapping(kernel)
    except NotImplementedError:
        return False
    return True


def kernel_structure(kernel):
    """What the theta mapping depends on, without the hyperparameter values: the
    tree of kernel types, every hyperparameter's bounds ("fixed" removes it from
    theta), the length-scale shapes and Matern's nu."""
    if isinstance(kernel, (kernels.Sum, kernels.Product)):
        return (type(kernel), kernel_structure(kernel.k1), kernel_structure(kernel.k2))
    if isinstance(kernel, kernels.Exponentiation):
        return (type(kernel), kernel_structure(kernel.kernel), kernel.exponent)
    attrs = vars(kernel)
    leaf = [type(kernel)]
    for name in sorted(attrs):
        v = attrs[name]
        if name.endswith("_bounds"):
            leaf.append((name, v if isinstance(v, str) else np.asarray(v).tobytes()))
        elif name == "length_scale":
            leaf.append((name, np.shape(v)))
        elif name == "nu":
            leaf.append((name, float(v)))
    return tuple(leaf)


class KernelDecomposition:
    """Cache of the decomposition of the last kernel seen (identity + theta),
    like CudaKernelAccelerator._update_kernel_if_needed (cuda_kernel_accelerators.py:146-164).
    The theta mapping is kept while the kernel's structure is unchanged (an
    optimizer only moves theta), so a new theta re-reads the parameter values
    only."""

    def __init__(self):
        self._kernel = None
        self._theta = None
        self._structure = None
        self.params = None
        self.mapping = None
        self.error = None

    def update(self, kernel):
        theta = np.asarray(kernel.theta)
        if self._kernel is kernel and self._theta is not None and self._theta.shape == theta.shape \
                and np.all(self._theta == theta):
            return self
        structure = kernel_structure(kernel)
        keep_mapping = self._kernel is kernel and self.mapping is not None and \
            self._theta is not None and self._theta.shape == theta.shape and \
            self._structure == structure
        self._kernel = kernel
        self._theta = copy.deepcopy(theta)
        self._structure = structure
        mapping = self.mapping if keep_mapping else None
        self.params, self.mapping, self.error = None, None, None
        try:
            self.params = get_kernel_params(kernel)
            self.mapping = mapping if mapping is not None else get_kernel_theta_mapping(kernel)
        except NotImplementedError as e:
            self.params, self.mapping, self.error = None, None, e
        return self

    @property
    def n_theta(self):
        return 0 if self._theta is None else int(self._theta.size)

    @property
    def accelerated(self):
        return self.params is not None
