"""FP64 CPU restatement of the reference time step: test infrastructure only."""
