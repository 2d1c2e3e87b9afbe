"""gRPC front end: ``ShadowVMAPI.RunPipelineCMD`` over libsvm.

The reference serves one RPC (com/svm.proto:157-159): a client sends a
``PipelineInfo``; the server resolves ``pre_pipeline_id[0]`` (source) and
``[1..]`` (ref objects) among its cached objects, runs the pipeline, caches the
output under ``PipelineInfo.id`` and answers ``Ret{success, info}``
(service/service.cu:25-63).  A missing dependency is answered with
``CANCELLED`` (service/service.cu:39,46).

Here the request bytes go to ``svm_pipeline_run`` unparsed: the C++ data plane
decodes the proto3 message, the device store is the object cache, and the
output stays resident under the pipeline id.  Where the reference aborts on a
failure (an assert), this server answers with a gRPC error status instead:

=====================  ========================
libsvm status           gRPC status
=====================  ========================
NOT_FOUND              CANCELLED (as reference)
INVALID                INVALID_ARGUMENT
UNSUPPORTED            UNIMPLEMENTED
BUSY / NO_MEMORY       RESOURCE_EXHAUSTED
STATE                  FAILED_PRECONDITION
DEVICE / JIT / other   INTERNAL
=====================  ========================

The service and method names are those of com/svm.proto (package ``ShadowVM``),
so the reference's clients (service/service.hpp:36-70, the Spark vRDD) talk to
it unchanged.  Python's grpcio is used because the C++ gRPC libraries are not
part of this toolchain; every RPC is one C call that releases the GIL.
"""
from __future__ import annotations

import argparse
import ctypes
import sys
from concurrent import futures

import grpc

from . import _lib
from .proto import Ret, serialize

SERVICE = "ShadowVM.ShadowVMAPI"
METHOD = "RunPipelineCMD"
PATH = f"/{SERVICE}/{METHOD}"

_CODES = {
    _lib.ERR_NOT_FOUND: grpc.StatusCode.CANCELLED,
    _lib.ERR_INVALID: grpc.StatusCode.INVALID_ARGUMENT,
    _lib.ERR_UNSUPPORTED: grpc.StatusCode.UNIMPLEMENTED,
    _lib.ERR_BUSY: grpc.StatusCode.RESOURCE_EXHAUSTED,
    _lib.ERR_NO_MEMORY: grpc.StatusCode.RESOURCE_EXHAUSTED,
    _lib.ERR_STATE: grpc.StatusCode.FAILED_PRECONDITION,
}


def _pipeline_id(data: bytes) -> str:
    """PipelineInfo.id (field 1) read from the wire bytes; '' if absent."""
    from .proto import PipelineInfo
    m = PipelineInfo()
    m.ParseFromString(data)
    return m.id


def run_pipeline_cmd(data: bytes, context) -> bytes:
    """One RunPipelineCMD: bytes in, serialized Ret out (or an error status)."""
    buf = (ctypes.c_uint8 * len(data)).from_buffer_copy(data)
    rc = _lib.lib().svm_pipeline_run(buf, len(data), None)
    if rc < 0:
        msg = _lib.lib().svm_last_error().decode(errors="replace")
        context.abort(_CODES.get(rc, grpc.StatusCode.INTERNAL), msg)
    return Ret(success=True, info="ShadowVM exec: " + _pipeline_id(data)).SerializeToString()


class ShadowVMService:
    """service/service.hpp:72-107: Launch(endpoint) / Wait(), over the device store."""

    def __init__(self, max_workers: int = 8):
        self.max_workers = max_workers
        self.server = None
        self.port = None

    def Launch(self, endpoint: str = "0.0.0.0:50051"):
        handler = grpc.method_handlers_generic_handler(SERVICE, {
            METHOD: grpc.unary_unary_rpc_method_handler(run_pipeline_cmd),  # raw bytes both ways
        })
        self.server = grpc.server(futures.ThreadPoolExecutor(max_workers=self.max_workers))
        self.server.add_generic_rpc_handlers((handler,))
        self.port = self.server.add_insecure_port(endpoint)
        if self.port == 0:
            raise RuntimeError(f"cannot listen on {endpoint}")
        self.server.start()
        return self.port

    def Wait(self):
        self.server.wait_for_termination()

    def Stop(self, grace: float = 0.0):
        if self.server is not None:
            self.server.stop(grace).wait()
            self.server = None


class ShadowVMClient:
    """service/service.hpp:36-70: a stub for RunPipelineCMD."""

    def __init__(self, channel):
        self.channel = channel
        self._call = channel.unary_unary(PATH, request_serializer=None, response_deserializer=None)

    @staticmethod
    def Create(endpoint: str = "localhost:50051") -> "ShadowVMClient":
        return ShadowVMClient(grpc.insecure_channel(endpoint))

    def RunPipelineCMD(self, pipeline, timeout: float | None = None):
        """Send a plan (dict or PipelineInfo bytes); returns the decoded Ret.
        Raises grpc.RpcError with the server's status on failure."""
        data = pipeline if isinstance(pipeline, (bytes, bytearray)) else serialize(pipeline)
        reply = Ret()
        reply.ParseFromString(self._call(bytes(data), timeout=timeout))
        return reply

    def close(self):
        self.channel.close()


def main(argv=None):
    """ShadowVM.cu's server role: initialise the device and serve until killed."""
    from . import pipeline as P
    ap = argparse.ArgumentParser(description="ShadowVM gRPC front end on libsvm")
    ap.add_argument("--endpoint", default="0.0.0.0:50051")
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--tuple-per-bulk", type=int, default=3_000_000)
    ap.add_argument("--workers", type=int, default=8)
    args = ap.parse_args(argv)
    P.init(device=args.device, tuple_per_bulk=args.tuple_per_bulk)
    svc = ShadowVMService(max_workers=args.workers)
    port = svc.Launch(args.endpoint)
    print(f"Server listening on port {port}", flush=True)
    try:
        svc.Wait()
    finally:
        P.shutdown()


if __name__ == "__main__":
    sys.exit(main())
