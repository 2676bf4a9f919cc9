"""Namespace of the ``brainpy.state_amd`` package (the MI355X network-step engine)."""
