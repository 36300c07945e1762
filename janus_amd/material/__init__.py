"""Materials (janus/material/__init__.py)."""
