"""The Python binding checks tensors before handing bare pointers to the C ABI
(CPU-only part: the helper itself)."""
import pytest

import mcfft_amd as mc


def test_check_tensor_rejects_host_tensors():
    torch = pytest.importorskip("torch")
    with pytest.raises(ValueError, match="not a device tensor"):
        mc._check_tensor(torch.zeros(16), 16, "input")
    with pytest.raises(ValueError, match="missing"):
        mc._check_tensor(None, 16, "input")
