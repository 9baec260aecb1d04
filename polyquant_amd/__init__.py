"""MI355X-native configuration-interaction hot path (sigma, Davidson, AO->MO, serving)."""
