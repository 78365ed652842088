"""FP64 CPU restatement of the TURBID RK3 substep (test infrastructure only)."""
