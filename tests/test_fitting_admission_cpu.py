"""Admission of card-fitting jobs (no GPU): the HBM budget and the target
pool refuse jobs that do not fit, and jobs beyond the fitting kernel's
limits are rejected with an error."""
import numpy as np
import pytest
import torch

from strns2cards_amd import fitting


def job(jid, res, quads):
    strands = np.zeros((3, 8, 3), np.float32)
    return fitting.FitJob(jid, res, quads, strands, np.zeros((24, 4), np.float32),
                          np.eye(4, dtype=np.float32),
                          np.zeros((2 * (quads + 1), 3), np.float32))


def server(capacity, budget):
    return fitting.CardFitServer(capacity, fitting.FitHyper(), torch.device("cpu"),
                                 budget_bytes=budget)


def test_job_bytes():
    j = job(0, 256, 32)
    assert j.words == 2048
    assert j.bytes == 4 * 2048 + 40 * 66 + fitting.SLOT_SCALAR_BYTES
    assert job(1, 100, 8).words == (100 * 100 + 31) // 32


def test_budget_refuses_jobs_that_do_not_fit():
    """The budget pays for the slot tables; the rest holds the target masks.
    With room for three 256^2 masks and one 64^2 mask, a fourth 256^2 job
    waits while the small one is admitted."""
    cap = 8
    table = cap * (40 * fitting.MAX_VERTS + fitting.SLOT_SCALAR_BYTES)
    big, small = job(0, 256, 32), job(1, 64, 8)
    srv = server(cap, table + 4 * (3 * 2048 + 128) + 8)
    assert srv.pool.words == 3 * 2048 + 128 + 2
    assert srv.can_admit(big)
    assert srv.can_admit(big, pending_words=[2048, 2048])
    assert not srv.can_admit(big, pending_words=[2048] * 3)
    assert srv.can_admit(small, pending_words=[2048] * 3)
    assert not srv.can_admit(small, pending_words=[2048] * 3 + [128])
    # a larger budget is capped at what the slots can use
    assert server(cap, 1e12).pool.words == cap * fitting.MAX_WORDS


def test_pool_bounds_admission():
    """The target pool holds `capacity` largest masks at most; with a pool of
    two 2048-word blocks a third 256^2 job waits for a release."""
    pool = fitting.WordPool(2 * 2048)
    a, b = pool.alloc(2048), pool.alloc(2048)
    assert (a, b) == (0, 2048)
    assert not pool.can_alloc(2048) and pool.alloc(2048) is None
    pool.release(a, 2048)
    assert pool.can_alloc(2048) and pool.alloc(2000) == a  # same block size class
    # small blocks come from their own size class or the bump pointer
    p2 = fitting.WordPool(1024)
    offs = [p2.alloc(128) for _ in range(8)]
    assert offs == list(range(0, 1024, 128)) and p2.alloc(128) is None
    p2.release(offs[3], 128)
    assert p2.alloc(100) == offs[3]


def test_slots_bound_admission():
    srv = server(2, 1e9)
    j = job(0, 64, 8)
    assert srv.can_admit(j, pending_words=[j.words])
    assert not srv.can_admit(j, pending_words=[j.words, j.words])


@pytest.mark.parametrize("res,quads", [(300, 8), (0, 8), (64, 0), (64, 33)])
def test_jobs_beyond_the_kernel_are_rejected(res, quads):
    srv = server(4, 1e9)
    with pytest.raises(ValueError):
        srv.can_admit(job(0, res, quads))
    with pytest.raises(ValueError):
        srv.admit([job(0, res, quads)])


def test_budget_too_small_for_the_tables():
    with pytest.raises(ValueError):
        server(1000, 1e5)
