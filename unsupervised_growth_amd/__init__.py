"""MI355X (gfx950) engine for the HVC network-growth trial of Sawndip/unsupervised_growth.

The hot path -- neuron integration, spike delivery through axonal-delay queues
and the STDP synapse-state machine -- runs as hand-written HIP kernels behind the
C ABI in include/hvc_engine.h (libhvc_engine.so, built by ``build.py``).
Python modules:
  engine    Engine: one network resident on one device
  network   Network arrays, parameter structs, synthetic networks
  configs   the benchmark / parity workloads
  opcount   algorithmic work per neuron-step
"""
__version__ = "0.1.0"
