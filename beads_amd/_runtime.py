"""One HIP runtime per process.

PyTorch ships its own ``libamdhip64.so.7`` (and HSA runtime) in ``torch/lib``;
``libbeads_amd.so`` names the same soname and finds ``/opt/rocm``'s copy by
its runpath. Whichever loads first serves both, and torch cannot initialise
its devices on top of the other one ("No HIP GPUs are available"). Loading
torch first, when it is installed, makes its runtime the shared one, which the
engine runs on equally well. Called before the engine library or the
``_native`` binding is loaded.
"""

from __future__ import annotations


def share_hip_runtime() -> None:
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
