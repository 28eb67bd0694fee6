"""Oracle package: CPU restatement of the reference algorithm (test infrastructure only)."""
