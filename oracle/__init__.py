"""FP64 CPU oracle (test infrastructure only; see uvlm_oracle.py)."""
