"""Oracle package: CPU FP64 restatement of the reference hot path (test infrastructure only)."""
