"""gRPC front end on the GPU: the reference's call sequence (dimension pipelines,
then the fact pipeline, each one RunPipelineCMD resolving its inputs among cached
objects) leaves the oracle's output objects (every dimension object and the result) and the
same output object as running the plans directly; a missing
dependency is answered with CANCELLED as in service/service.cu:39,46."""
import grpc
import numpy as np
import pytest

from shadowvm_amd import pipeline as P
from shadowvm_amd import service as S
from shadowvm_amd import ssb

from .helpers import assert_same_columns, host_tables, oracle_query

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def server(engine):
    svc = S.ShadowVMService(max_workers=4)
    port = svc.Launch("127.0.0.1:0")
    client = S.ShadowVMClient.Create(f"127.0.0.1:{port}")
    yield client
    client.close()
    svc.Stop()


def test_star_join_through_grpc_equals_direct(server):
    wl = ssb.four_way(seed=41, fact_rows=2_000_000, n_cust=200_000, n_supp=20_000, n_part=100_000,
                      prefix="gr")
    ssb.load_device(wl)
    for pt, src, refs in wl.dims:
        ret = server.RunPipelineCMD(pt.plan(src, refs), timeout=120)
        assert ret.success and ret.info == "ShadowVM exec: " + pt.pipe_info["id"]
    fq, src, refs = wl.fact
    ret = server.RunPipelineCMD(fq.plan(src, refs), timeout=120)
    assert ret.success
    fid = fq.pipe_info["id"]
    rows, ncols = P.table_info(fid)
    got = [P.table_column(fid, c) for c in range(ncols)]
    # the object the gRPC call cached equals the oracle's answer to the same plans
    want, want_dims = oracle_query(wl, host_tables(wl))
    assert rows == want.rows and rows > 0
    assert_same_columns(got, want.columns)
    for pt, _, _ in wl.dims:
        pid = pt.pipe_info["id"]
        n, nc = P.table_info(pid)
        w = want_dims[pid]
        assert n == w.rows
        assert_same_columns([P.table_column(pid, c) for c in range(len(w.columns))], w.columns)
    direct = fq.Run(src, refs)  # the same plan run in-process (re-caches the same id)
    assert rows == direct.rows
    for g, d in zip(got, direct.columns):
        assert np.array_equal(np.asarray(g), np.asarray(d))


def test_missing_dependency_is_cancelled(server):
    wl = ssb.single_join(seed=2, fact_rows=1000, prefix="gm")
    fq, src, refs = wl.fact
    with pytest.raises(grpc.RpcError) as e:
        server.RunPipelineCMD(fq.plan("no_such_object", refs), timeout=60)
    assert e.value.code() == grpc.StatusCode.CANCELLED
