"""MI355X-native solve backend for the ohmnivore circuit simulator."""
