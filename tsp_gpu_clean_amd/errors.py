"""Exceptions of the package (mirror the reference's src/data_models/exceptions.py names)."""


class GPUUnavailableError(RuntimeError):
  """No HIP device or no built libtspmi.so; GPU strategies never fall back to CPU.

  Mirrors the reference's ``GPUUnavailableError`` raised by ``require_gpu``
  (src/protocols/backend.py:235-244, used at two_opt.py:746-754).
  """


class VRAMInsufficientError(MemoryError):
  """A dense allocation would not fit in free device memory (problem_context.py:367-374)."""
