"""Spike-train comparison for the precision forms.

The FP32 form is judged against the FP64 oracle by the spike sets and times it
produces: a spike of one train and a spike of the other train of the same
neuron are the same event when they lie within ``window`` ms of each other
(paired in time order, nearest first).  Reported: how many spikes each side has,
how many could not be paired (the spike-set mismatch), and the largest |dt|
over the paired spikes.
"""
import numpy as np


def _pair(a, b, window):
    """Greedy in-order pairing of two sorted spike-time arrays."""
    i = j = 0
    dts = []
    while i < len(a) and j < len(b):
        d = b[j] - a[i]
        if abs(d) <= window:
            dts.append(d)
            i += 1
            j += 1
        elif a[i] < b[j]:
            i += 1
        else:
            j += 1
    return dts


def compare_trains(ref, test, window=1.0, step=0.02):
    """ref, test: lists (one entry per neuron) of spike-time arrays in ms;
    step: the time step, for the count of paired spikes within one step."""
    if len(ref) != len(test):
        raise ValueError("trains cover different neuron counts: %d vs %d" % (len(ref), len(test)))
    n_ref = n_test = paired = 0
    worst = 0.0
    neurons_differ = 0
    within = 0
    for a, b in zip(ref, test):
        a = np.asarray(a, float)
        b = np.asarray(b, float)
        n_ref += len(a)
        n_test += len(b)
        dts = _pair(a, b, window)
        paired += len(dts)
        if dts:
            ad = np.abs(dts)
            worst = max(worst, float(np.max(ad)))
            within += int(np.count_nonzero(ad <= step + 1e-9))
        if len(dts) != len(a) or len(dts) != len(b):
            neurons_differ += 1
    return {
        "spikes_ref": n_ref,
        "spikes_test": n_test,
        "paired": paired,
        "unpaired": (n_ref - paired) + (n_test - paired),
        "mismatch_frac": ((n_ref - paired) + (n_test - paired)) / max(1, n_ref + n_test),
        "max_abs_dt_ms": worst,
        "paired_within_one_step": within,
        "neurons_differ": neurons_differ,
    }


def compare_results(ref, test, window=1.0, step=0.02):
    """Per kind (soma, dend, interneuron) comparison of two TrialResults."""
    return {k: compare_trains(getattr(ref, k), getattr(test, k), window, step) for k in ("soma", "dend", "interneuron")}
