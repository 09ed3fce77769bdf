"""MI355X-native GRPP integral engine (libgrpp drop-in)."""
