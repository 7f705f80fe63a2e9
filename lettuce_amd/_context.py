"""Execution context: device, arithmetic dtype and population storage.

Mirrors lettuce's ``Context`` (lettuce/_context.py:9-122): the same
constructor arguments, defaults and tensor factories.  Differences:

* the step always runs through the HIP kernels, so a context that is to run
  a simulation must live on a ROCm device (``use_native`` may not be False
  there);
* ``storage`` selects how populations are held in HBM: ``"f32"``/``"f64"``
  (the dtype itself) or ``"f16_shifted"``/``"bf16_shifted"`` (f_q - w_q rho0
  in 16 bits, FP32 arithmetic).  lettuce has no such option.
* ``layout`` selects the memory order behind the [Q, *res] view of the
  populations: ``"soa"`` (lettuce's contiguous [Q, *res]) or ``"rows"``
  (each row of the last axis stored for all Q populations together,
  [*res[:-1], Q, res[-1]]; periodic flows without boundaries only).
* ``rounding`` selects how the step stores ``"bf16_shifted"`` populations:
  ``"nearest"`` (round to nearest even, the default) or ``"stochastic"``
  (the dropped bits replaced by random bits before truncation, unbiased on
  average; the random bits follow from ``rounding_seed``, the step and the
  element, so runs are reproducible).  Other storage formats round to
  nearest.
"""
from typing import List, Optional, Union

import numpy as np
import torch

__all__ = ["Context"]

_STORAGE_DTYPES = {"f32": torch.float32, "f64": torch.float64,
                   "f16_shifted": torch.float16,
                   "bf16_shifted": torch.bfloat16}


class Context:
    def __init__(self, device: Optional[Union[torch.device, str]] = None,
                 dtype: Optional[torch.dtype] = None,
                 use_native: Optional[bool] = None,
                 storage: Optional[str] = None, layout: str = "soa",
                 rounding: str = "nearest", rounding_seed: int = 0x2545F491):
        cuda = torch.cuda.is_available()
        if device is None:
            if use_native is None:
                use_native = cuda
            device = torch.device("cuda:0" if cuda else "cpu")
        device = torch.device(device)
        if device.type not in ("cuda", "cpu"):
            raise ValueError(f"lettuce_amd runs on ROCm GPUs (torch 'cuda' "
                             f"devices); {device} is not supported")
        if device.type == "cuda":
            if not cuda:
                raise RuntimeError("a ROCm device was requested but none is "
                                   "available")
            if use_native is False:
                raise ValueError(
                    "lettuce_amd has no torch step: use_native=False is not "
                    "available on a GPU device")
            use_native = True
        else:
            if use_native:
                raise ValueError("the native step needs a ROCm device")
            use_native = False

        dtype = dtype or torch.float32
        if dtype not in (torch.float32, torch.float64):
            raise ValueError(
                f"arithmetic dtype must be float32 or float64, got {dtype}; "
                f"16-bit populations are selected with storage="
                f"'f16_shifted' or 'bf16_shifted' (FP32 arithmetic)")
        if storage is None:
            storage = "f64" if dtype == torch.float64 else "f32"
        if storage not in _STORAGE_DTYPES:
            raise ValueError(f"unknown storage {storage!r}; expected one of "
                             f"{sorted(_STORAGE_DTYPES)}")
        if dtype == torch.float64 and storage != "f64":
            raise ValueError("float64 arithmetic needs storage='f64'")
        if dtype == torch.float32 and storage == "f64":
            raise ValueError("storage='f64' needs dtype=torch.float64")

        if layout not in ("soa", "rows"):
            raise ValueError(f"unknown layout {layout!r}; expected 'soa' or 'rows'")
        if rounding not in ("nearest", "stochastic"):
            raise ValueError(f"unknown rounding {rounding!r}; expected 'nearest' or "
                             f"'stochastic'")
        if rounding == "stochastic" and storage != "bf16_shifted":
            raise ValueError("stochastic rounding applies to storage='bf16_shifted' only")
        seed = int(rounding_seed) & 0xFFFFFFFF
        if rounding == "stochastic" and seed == 0:
            raise ValueError("rounding_seed must be non-zero")

        self.device = device
        self.dtype = dtype
        self.use_native = use_native
        self.storage = storage
        self.layout = layout
        self.rounding = rounding
        # 0 tells the kernels to round to nearest
        self.rounding_seed = seed if rounding == "stochastic" else 0

    def population_tensor(self, q: int, resolution) -> torch.Tensor:
        """Uninitialised populations [q, *resolution] in the storage dtype,
        laid out as `layout` says."""
        res = list(resolution)
        if self.layout == "soa":
            return torch.empty([q] + res, device=self.device,
                               dtype=self.storage_dtype)
        raw = torch.empty(res[:-1] + [q, res[-1]], device=self.device,
                          dtype=self.storage_dtype)
        d = len(res)
        return raw.permute([d - 1] + list(range(d - 1)) + [d])

    @property
    def storage_dtype(self) -> torch.dtype:
        return _STORAGE_DTYPES[self.storage]

    @property
    def shifted(self) -> bool:
        return self.storage.endswith("_shifted")

    def synchronize(self):
        if self.device.type == "cuda":
            torch.cuda.synchronize(self.device)

    def empty_tensor(self, size: Union[List[int], torch.Size], *args,
                     dtype=None, **kwargs) -> torch.Tensor:
        return torch.empty(size, *args, **kwargs, device=self.device,
                           dtype=(dtype or self.dtype))

    def zero_tensor(self, size, *args, dtype=None, **kwargs) -> torch.Tensor:
        return torch.zeros(size, *args, **kwargs, device=self.device,
                           dtype=(dtype or self.dtype))

    def full_tensor(self, size, value, *args, dtype=None, **kwargs
                    ) -> torch.Tensor:
        return torch.full(size, value, *args, **kwargs, device=self.device,
                          dtype=(dtype or self.dtype))

    def one_tensor(self, size, *args, dtype=None, **kwargs) -> torch.Tensor:
        return torch.ones(size, *args, **kwargs, device=self.device,
                          dtype=(dtype or self.dtype))

    def convert_to_tensor(self, array, *args,
                          dtype: Optional[torch.dtype] = None, **kwargs
                          ) -> torch.Tensor:
        if dtype is None:
            bool_like = (bool, torch.bool, torch.uint8, np.uint8,
                         np.dtype(bool), np.dtype(np.uint8))
            if hasattr(array, "dtype") and array.dtype in bool_like:
                dtype = torch.uint8
            else:
                dtype = self.dtype
        if isinstance(array, torch.Tensor):
            return array.to(*args, **kwargs, device=self.device, dtype=dtype)
        array = np.ascontiguousarray(np.array(array))
        return torch.from_numpy(array).to(*args, **kwargs, device=self.device,
                                          dtype=dtype)

    @staticmethod
    def convert_to_ndarray(tensor) -> np.ndarray:
        if isinstance(tensor, torch.Tensor):
            return tensor.detach().cpu().numpy()
        return np.array(tensor)
