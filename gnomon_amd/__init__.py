"""gnomon_amd -- MI355X (gfx950) backend for gnomon's polygenic-scoring hot path.

The compute lives in ``libgnomon_hip.so`` (HIP kernels + C ABI,
include/gnomon_hip.h).  This package is the host-side mirror of the reference
interface around it:

* ``prep``      -- weight canonicalisation into the reference CSR
                   (score/prepare.rs:197-208, 632-650) and the .sscore columns;
* ``pipeline``  -- ``try_run_hip`` / ``ScoringContext``, the contract of
                   score/cuda_backend.rs:604 ``try_run_cuda``;
* ``synth``     -- synthetic workloads (device-side genotype generator);
* ``autograd``  -- ``torch.autograd.Function`` over the forward/adjoint entry
                   points (imported lazily: needs torch).
"""
from ._lib import GnmError, lib  # noqa: F401
from .prep import ScoreCsr, csr_from_dense, csr_from_records, sscore_columns  # noqa: F401
from .pipeline import ScoringContext, try_run_hip  # noqa: F401
