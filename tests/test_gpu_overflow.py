"""Staging overflow is reported, not written out of bounds.

Each synapse delivers at most one event per step, so a tile's staging region
sized by its targets' incoming synapses never overflows; when the staging would
exceed the memory budget (`bucket_mem_frac`), every region is scaled down by the
same factor, and a step whose events do not fit stops the run with an error
(status bit 1) after the gathers have written what fits.  Here the budget is
forced from the host to 0.05 % of the synapses, so that the CUBA network's
average activity (about 1.5 events per neuron and step) overflows; the kernels bound every write by its region.
"""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [{}, {"tile": 1024}], ids=["separate", "fused-tile1024"])
def test_staging_overflow_is_reported(shape):
    import torch

    from brian2_amd import synth
    from brian2_amd._lib import HipBrianError

    N, K = 5000, 1000
    free, _ = torch.cuda.mem_get_info()
    # regions of 0.05 % of each tile's in-degree (0.5 events per target): 3 buffers of
    # 12-byte records
    frac = (N * K // 2000) * 3 * 12 / free
    print(f"staging overflow, N={N}, K={K}, shape {shape}, bucket_mem_frac {frac:.3g}", flush=True)
    c = synth.cuba_synw_network(N=N, K=K, network_options=dict(shape, bucket_mem_frac=frac))
    with pytest.raises(HipBrianError, match="staging overflow"):
        c["net"].run(3000)


@pytest.mark.parametrize("shape", [{}, {"tile": 1024}], ids=["separate", "fused-tile1024"])
def test_full_staging_runs_clean(shape):
    """The same network with the default budget: no overflow, no error."""
    from brian2_amd import synth

    c = synth.cuba_synw_network(N=5000, K=1000, network_options=dict(shape))
    c["net"].run(3000)
    assert c["net"].n_events > 0
