"""Linearisations beyond EK1 (reference ``pof/linearization``)."""
