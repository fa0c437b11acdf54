"""Exceptions raised for the C ABI status codes of libflowfilter_amd.

The reference's C++ classes throw std::invalid_argument, std::logic_error or
a bare std::exception (src/gpu/image.cu:117, src/gpu/flowfilter.cu:53,88,
src/gpu/update.cu:126-129); its Cython bindings declare no ``except +``, so
there those terminate the interpreter. Here each maps to a Python exception:

    FF_ERR_INVALID_ARGUMENT (std::invalid_argument) -> InvalidArgument (ValueError)
    FF_ERR_LOGIC            (std::logic_error)      -> LogicError (RuntimeError)
    FF_ERR_RUNTIME          (std::exception)        -> FilterError (RuntimeError)
    FF_ERR_HIP              (CUDA error -> exit())  -> HipError (RuntimeError)
    FF_ERR_BAD_ALLOC                                -> MemoryError
    FF_ERR_NULL_HANDLE                              -> InvalidArgument
    FF_ERR_WRONG_TYPE                               -> TypeError
"""

FF_OK = 0
FF_ERR_INVALID_ARGUMENT = 1
FF_ERR_LOGIC = 2
FF_ERR_RUNTIME = 3
FF_ERR_HIP = 4
FF_ERR_BAD_ALLOC = 5
FF_ERR_NULL_HANDLE = 6
FF_ERR_WRONG_TYPE = 7


class InvalidArgument(ValueError):
    pass


class LogicError(RuntimeError):
    pass


class FilterError(RuntimeError):
    pass


class HipError(RuntimeError):
    pass


_MAP = {
    FF_ERR_INVALID_ARGUMENT: InvalidArgument,
    FF_ERR_LOGIC: LogicError,
    FF_ERR_RUNTIME: FilterError,
    FF_ERR_HIP: HipError,
    FF_ERR_BAD_ALLOC: MemoryError,
    FF_ERR_NULL_HANDLE: InvalidArgument,
    FF_ERR_WRONG_TYPE: TypeError,
}


def raise_for_status(status, message):
    raise _MAP.get(status, FilterError)(message)
