"""ORACLE -- test infrastructure only: the CPU re-statement of the reference pipeline used as the checker."""
