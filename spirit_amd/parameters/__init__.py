"""Method parameters (core/python/spirit/parameters)."""
