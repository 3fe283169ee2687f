"""NumpyNngpr: the reference's numpy facade (numpy_nngpr.py:15-90) on the HIP engine.

Same constructor and numpy inputs/outputs as the reference; the arithmetic runs
on HIP device 0 (arrays are copied in and out).  `num_workers` is accepted for
compatibility: there is one device worker.
"""
from numbers import Integral

from sklearn.utils._param_validation import Interval

from .base_nngpr import HipWorker, Nngpr


class NumpyNngpr(Nngpr):
    _parameter_constraints: dict = {
        **Nngpr._parameter_constraints,
        "num_workers": [Interval(Integral, 1, None, closed="left"), None],
    }

    def __init__(self, kernel=None, *, num_workers=None, alpha=1e-10, optimizer="fmin_l_bfgs_b",
                 n_restarts_optimizer=0, normalize_y=False, copy_X_train=True, random_state=None,
                 num_nn=None, nn_type=None, batch_size=None, allow_downcast_f32=None,
                 distribute_method=None, mp_start_method=None, random_gen_batch_byte_size=None):
        self.num_workers = num_workers
        super().__init__([HipWorker("cuda:0")], kernel=kernel, alpha=alpha, optimizer=optimizer,
                         n_restarts_optimizer=n_restarts_optimizer, normalize_y=normalize_y,
                         copy_X_train=copy_X_train, random_state=random_state, num_nn=num_nn,
                         nn_type=nn_type, batch_size=batch_size,
                         allow_downcast_f32=allow_downcast_f32,
                         distribute_method=distribute_method, mp_start_method=mp_start_method,
                         random_gen_batch_byte_size=random_gen_batch_byte_size)
