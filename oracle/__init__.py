"""The CPU oracle (test infrastructure; see loader.py)."""
