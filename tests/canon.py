"""Re-export of the package's host gauge helpers for the tests."""
from qutree_amd.gauge import canonicalize_all, canonicalize_path, node_view, push_up  # noqa: F401
