"""TensorFox for MI355X: the dense CPD path of Tensor Fox (felipebottega/Tensor-Fox v1.2.1,
modules/TensorFox/__init__.py) on hand-written gfx950 HIP kernels.

    import TensorFox as tfx
    factors, output = tfx.cpd(T, R, options)
    factors, output = tfx.cpd([data, idxs, dims], R, options)   # sparse COO input
    loss = tfx.cpd_loss(T, A, B, C)      # differentiable 0.5 ||T - [[A, B, C]]||^2
    results = tfx.cpd_batched(tensors, ranks, options, seeds)   # many small CPDs at once
"""

__version__ = '1.2.1+mi355x'


def version():
    return (1, 2, 1)


from .auxiliar import (Output, consistency, gen_rand_tensor, make_lowrank, make_options,  # noqa: E402
                       sort_dims, unsort_dims, cpd2tens, gen_rand_sparse_tensor, sparse2dense,
                       sparse2txt)
from .tensorfox import cpd, tricpd  # noqa: E402
from .loss import CPDLoss, cpd_loss  # noqa: E402
from .batched import BatchedCPD, autotune_batched, cpd_batched  # noqa: E402
from .drivers import foxit, rank, stats  # noqa: E402
