"""Drop-in for the reference's tensor_class package."""
