"""FP64 CPU oracle (test infrastructure): see refcpu.h."""
