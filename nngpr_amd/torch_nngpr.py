"""TorchNngpr: the reference's torch facade (torch_nngpr.py:17-106) on the HIP engine."""
import torch

from .base_nngpr import HipWorker, Nngpr


class TorchNngpr(Nngpr):
    """NNGP regressor whose computations run as HIP kernels on one device.

    Parameters
    ----------
    torch_devices : str or list[str]
        'cuda', 'cuda:N' or a list of them.  The engine runs on one GPU: the first
        HIP device of the list is used.  CPU devices are rejected.
    Other parameters: as the reference's Nngpr (kernel, alpha, optimizer,
    n_restarts_optimizer, normalize_y, copy_X_train, random_state, num_nn,
    nn_type, batch_size, allow_downcast_f32, distribute_method, mp_start_method,
    random_gen_batch_byte_size, distribute_environ).
    """

    _parameter_constraints: dict = {**Nngpr._parameter_constraints}

    def __init__(self, torch_devices, kernel=None, *, alpha=1e-10, optimizer="fmin_l_bfgs_b",
                 n_restarts_optimizer=0, normalize_y=False, copy_X_train=True, random_state=None,
                 num_nn=None, nn_type=None, batch_size=None, allow_downcast_f32=None,
                 distribute_method=None, mp_start_method=None, random_gen_batch_byte_size=None,
                 distribute_environ=None):
        kernel = self.get_default_kernel() if kernel is None else kernel
        if isinstance(torch_devices, str):
            torch_devices = ["cuda:0"] if torch_devices == "cuda" else [torch_devices]
        self.torch_devices = torch_devices
        hip = [d for d in torch_devices if torch.device(d).type == "cuda"]
        if not hip:
            raise ValueError("TorchNngpr on the HIP engine needs a HIP device ('cuda[:N]'), "
                             "got %r" % (torch_devices,))
        workers = [HipWorker(hip[0])]
        super().__init__(workers, kernel=kernel, alpha=alpha, optimizer=optimizer,
                         n_restarts_optimizer=n_restarts_optimizer, normalize_y=normalize_y,
                         copy_X_train=copy_X_train, random_state=random_state, num_nn=num_nn,
                         nn_type=nn_type, batch_size=batch_size,
                         allow_downcast_f32=allow_downcast_f32,
                         distribute_method=distribute_method, mp_start_method=mp_start_method,
                         random_gen_batch_byte_size=random_gen_batch_byte_size,
                         distribute_environ=distribute_environ)
