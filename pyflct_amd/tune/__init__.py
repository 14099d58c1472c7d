"""Offline tuning of build-time kernel parameters (see sweep_slots.py)."""
