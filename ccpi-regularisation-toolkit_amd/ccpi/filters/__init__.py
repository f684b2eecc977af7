"""MI355X regularisers with the reference toolkit's ``ccpi.filters`` surface."""
