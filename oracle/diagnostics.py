"""ESS and R-hat in numpy float64 (oracle).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates geomjax/diagnostics.py:25-75 (potential_scale_reduction) and
:78-209 (effective_sample_size: FFT autocovariance, Geyer initial positive and
initial monotone sequences, tau floored at 1/log10(M N)).
"""
import numpy as np
from scipy.fftpack import next_fast_len


def potential_scale_reduction(input_array, chain_axis=0, sample_axis=1):
    x = np.asarray(input_array, dtype=np.float64)
    assert x.shape[chain_axis] > 1, "potential_scale_reduction as implemented only works for two or more chains."
    num_samples = x.shape[sample_axis]
    per_chain_mean = x.mean(axis=sample_axis, keepdims=True)
    per_chain_var = x.var(axis=sample_axis, ddof=1, keepdims=True)
    between = num_samples * per_chain_mean.var(axis=chain_axis, ddof=1, keepdims=True)
    within = per_chain_var.mean(axis=chain_axis, keepdims=True)
    rhat = np.sqrt((between / within + num_samples - 1) / num_samples)
    return rhat.squeeze()


def effective_sample_size(input_array, chain_axis=0, sample_axis=1):
    x = np.asarray(input_array, dtype=np.float64)
    sample_axis = sample_axis if sample_axis >= 0 else x.ndim + sample_axis
    num_chains = x.shape[chain_axis]
    num_samples = x.shape[sample_axis]
    assert num_chains > 1, "effective_sample_size as implemented only works for two or more chains."

    mean_across_chain = x.mean(axis=sample_axis, keepdims=True)
    centered = x - mean_across_chain
    m = next_fast_len(2 * num_samples)
    f = np.fft.rfft(centered, n=m, axis=sample_axis)
    f = f * np.conjugate(f)
    autocov = np.fft.irfft(f, n=m, axis=sample_axis)
    autocov = np.take(autocov, np.arange(num_samples), axis=sample_axis) / num_samples
    mean_autocov_var = autocov.mean(chain_axis, keepdims=True)
    mean_var0 = np.take(mean_autocov_var, np.array([0]), axis=sample_axis) * num_samples / (num_samples - 1.0)
    weighted_var = mean_var0 * (num_samples - 1.0) / num_samples
    weighted_var = weighted_var + mean_across_chain.var(axis=chain_axis, ddof=1, keepdims=True)

    num_samples_even = num_samples - num_samples % 2
    tp1 = np.take(mean_autocov_var, np.arange(1, num_samples_even), axis=sample_axis)
    rho_hat = np.concatenate([np.ones_like(mean_var0), 1.0 - (mean_var0 - tp1) / weighted_var], axis=sample_axis)
    rho_hat = np.moveaxis(rho_hat, sample_axis, 0)
    rho_even = rho_hat[0::2]
    rho_odd = rho_hat[1::2]

    # Geyer's initial positive sequence (:147-182)
    mask0 = (rho_even + rho_odd) > 0.0
    carry = np.ones_like(mask0[0])
    max_t = np.zeros(mask0[0].shape, dtype=np.int64)
    mask = np.empty_like(mask0)
    for t in range(mask0.shape[0]):
        carry = carry & mask0[t]
        max_t = np.where(carry, t, max_t)
        mask[t] = carry
    idx = np.indices(max_t.shape)
    indices = tuple([max_t + 1] + [idx[i] for i in range(max_t.ndim)])
    rho_odd = np.where(mask, rho_odd, 0.0)
    # jnp's .at[].set with an out-of-range index drops the update; mirror that.
    mask_even = mask.copy()
    inb = indices[0] < rho_even.shape[0]
    sel = tuple(ix[inb] for ix in indices)
    mask_even[sel] = rho_even[sel] > 0
    rho_even = np.where(mask_even, rho_even, 0.0)

    # Geyer's initial monotone sequence (:184-196)
    rho_sum = rho_even + rho_odd
    prev = rho_sum[0]
    update_mask = np.empty_like(rho_sum, dtype=bool)
    update_value = np.empty_like(rho_sum)
    for t in range(rho_sum.shape[0]):
        um = rho_sum[t] > prev
        nv = np.where(um, prev, rho_sum[t])
        update_mask[t] = um
        update_value[t] = nv
        prev = nv
    rho_even_final = np.where(update_mask, update_value / 2.0, rho_even)
    rho_odd_final = np.where(update_mask, update_value / 2.0, rho_odd)

    ess_raw = num_chains * num_samples
    # jnp indexing clamps an out-of-range gather index; mirror that.
    gidx = (np.minimum(indices[0], rho_even_final.shape[0] - 1),) + indices[1:]
    tau_hat = -1.0 + 2.0 * np.sum(rho_even_final + rho_odd_final, axis=0) - rho_even_final[gidx]
    tau_hat = np.maximum(tau_hat, 1 / np.log10(ess_raw))
    return (ess_raw / tau_hat).squeeze()
