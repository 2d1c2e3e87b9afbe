"""gRPC front end without a GPU: the service answers on the reference's method
path (com/svm.proto:157-159) with raw PipelineInfo bytes, and libsvm statuses map
to gRPC codes (a run before svm_init is FAILED_PRECONDITION; the reference's
CANCELLED for a missing dependency is covered on the GPU)."""
import grpc
import pytest

from shadowvm_amd import service as S
from shadowvm_amd import ssb
from shadowvm_amd.proto import PipelineInfo, Ret, serialize


@pytest.fixture
def server():
    svc = S.ShadowVMService(max_workers=2)
    port = svc.Launch("127.0.0.1:0")
    client = S.ShadowVMClient.Create(f"127.0.0.1:{port}")
    yield client
    client.close()
    svc.Stop()


def test_method_path_and_messages():
    assert S.PATH == "/ShadowVM.ShadowVMAPI/RunPipelineCMD"
    r = Ret(success=True, info="ShadowVM exec: q1")
    back = Ret()
    back.ParseFromString(r.SerializeToString())
    assert back.success and back.info == "ShadowVM exec: q1"
    data = serialize(ssb.single_join(seed=1, fact_rows=10).fact[0].plan("src", ["d"]))
    assert S._pipeline_id(data) == "sj_query"


def test_run_before_init_is_failed_precondition(server):
    plan = ssb.single_join(seed=1, fact_rows=10).fact[0].plan("x", [])
    with pytest.raises(grpc.RpcError) as e:
        server.RunPipelineCMD(plan, timeout=30)
    assert e.value.code() == grpc.StatusCode.FAILED_PRECONDITION
    assert "svm_init" in e.value.details()


def test_unknown_method_is_unimplemented(server):
    call = server.channel.unary_unary("/ShadowVM.ShadowVMAPI/Nope", request_serializer=None,
                                      response_deserializer=None)
    with pytest.raises(grpc.RpcError) as e:
        call(PipelineInfo(id="x").SerializeToString(), timeout=30)
    assert e.value.code() == grpc.StatusCode.UNIMPLEMENTED
