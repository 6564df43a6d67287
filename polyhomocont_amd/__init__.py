"""MI355X path tracker for polyhomocont's total-degree homotopy."""
