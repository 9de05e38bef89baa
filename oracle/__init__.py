"""TEST INFRASTRUCTURE: the CPU oracle package (checker only, never shipped)."""
