"""Device/dtype defaults of the reference (utils.py:7-41)."""

from typing import Optional, Tuple, Union

import torch


def default(value, default_value):
    """Return default_value if value is None (utils.py:7-18)."""
    return value if value is not None else default_value


def format_device_dtype(
    device: Optional[Union[torch.device, str]] = None, dtype: Optional[torch.dtype] = None
) -> Tuple[torch.device, torch.dtype]:
    """CPU / float32 by default; "cuda" becomes cuda:0 (utils.py:20-41)."""
    if device is None:
        device = torch.device("cpu")
    elif isinstance(device, str):
        device = torch.device(device)
        if device.index is None and device.type != "cpu":
            device = torch.device(device.type, 0)
    dtype = default(dtype, torch.float32)
    return device, dtype
